/*
 * sph_capi.h -- C ABI of the MI355X (gfx950) delta-SPH dam-break step.
 *
 * Two layers are exported:
 *
 *  1. The reference call surface, one entry point per reference host wrapper,
 *     with the same argument roles, the same device-pointer ownership (the
 *     caller allocates every array) and the same array layouts (AoS float[3N]
 *     for positions, velocities and accelerations, int[N] for types).  The
 *     process-wide parameter block set by sph_set_params() replaces the
 *     reference's `__constant__ SimParams par`.  A caller that drove the
 *     reference step (opt_src_v0/sph/sph_run.cu:269-309) can call these in the
 *     same order.
 *
 *  2. A simulation context (sph_sim_*) that owns sorted SoA state with FP64
 *     positions and runs whole predictor-corrector steps on one HIP stream,
 *     optionally replayed from a captured hipGraph.  This is the fast path.
 *
 * Every function returns an int status: SPH_OK (0) on success, a negative
 * SPH_E* code on a usage error, or a positive hipError_t value when the HIP
 * runtime failed.  sph_last_error() returns a message for the last failure
 * on the calling thread.  Nothing here prints or pauses.
 *
 * 2-D convention: the plane is (x, z); gravity acts along z; y is 0 for every
 * particle and the grid has one cell along y (gridydim == 1).
 *
 * Particle types follow the reference (src/sph/sph_run.cu:36):
 * 0 wall (dummy particle), 1 fluid, anything else is treated as non-fluid.
 */
#ifndef SPH_CAPI_H
#define SPH_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_OK 0
#define SPH_E_INVAL (-1)     /* bad argument (null pointer, size out of range) */
#define SPH_E_NOPARAMS (-2)  /* sph_set_params was never called */
#define SPH_E_STATE (-3)     /* call sequence error (e.g. graph replay before capture) */
#define SPH_E_NODEV (-4)     /* no HIP device visible */

#define SPH_EOS_LINEAR 0     /* p = cs^2 (rho - rho0)   (src/sph/sph_run.cu:467) */
#define SPH_EOS_TAIT 1       /* p = rho0 cs^2 / gamma ((rho/rho0)^gamma - 1) */

/*
 * Parameter block.  Replaces SimParams (src/header/data_struct.cuh:4-44) and
 * its copy to `__constant__ par` (src/sph/sph_run.cu:105).  Only fields the
 * step reads are kept; derived constants are filled by the builders below.
 */
typedef struct sph_params {
    /* escape box: a fluid particle outside [min,max] on any axis is parked
       (src/sph/sph_run.cu:437-447; the reference hard-codes 0..3.23, 0..1, 0..1) */
    double domain_min[3];
    double domain_max[3];

    int dim;             /* 2 or 3 */
    int numParticles;    /* N, fluid + wall */
    int numFluid;        /* informational */
    int gridxdim, gridydim, gridzdim;
    int numCells;        /* gridxdim * gridydim * gridzdim */
    int hash_max;        /* reference value X*Y*Z + X*Y + X (src/sph/sph_run.cu:92); unused by the kernels */
    int eos;             /* SPH_EOS_LINEAR or SPH_EOS_TAIT */
    int park_type;       /* type written to a parked particle (1 keeps it fluid as in src/v0/v1/v3) */

    float gridxmin, gridymin, gridzmin;
    float gridSize;      /* cell edge */
    float particleR;     /* lattice spacing dx */
    float h, kh, eta, delta, afa;
    float adh;           /* Wendland C2 normalisation: 21/(16 pi h^3) in 3-D, 7/(4 pi h^2) in 2-D */
    float particleMass;  /* dx^dim * rho0 */
    float restDensity;   /* rho0 */
    float cs;            /* speed of sound c0 */
    float gravity;       /* along z, negative */
    float timeStep;      /* full step dt; each half-step advances dt/2 */
    float gamma;         /* Tait exponent */
    float tait_B;        /* rho0 cs^2 / gamma */
    float damper;        /* 1: acoustic damper dofv term on (ns.cu:212), 0: off */
    float rr_cut;        /* largest float rr with sqrtf(rr)/h <= 2 in FP32 (same pair set as the reference test q <= 2) */
    float park_pos[3];   /* where an escaped particle is parked (src: (0, 0.5, -0.08)) */
    float pad_;
} sph_params;

/* Dam-break case description (no reference counterpart: the reference reads
   its initial state from .dat files that are not shipped, src/io/io.cu:6-111). */
typedef struct sph_dambreak_cfg {
    int dim;             /* 2 or 3 */
    int eos;             /* SPH_EOS_LINEAR / SPH_EOS_TAIT */
    int wall_layers;     /* dummy wall layers (3) */
    int pad_;
    double dx;           /* lattice spacing */
    double tank[3];      /* tank extent x, y, z (y ignored in 2-D) */
    double column[3];    /* water column extent x, y, z at the x=0, y=0 corner */
    double h_over_dx;    /* 1.3 */
    double rho0;         /* 1000 */
    double g;            /* -9.81 */
    double c0;           /* <= 0: 10 sqrt(2 |g| H) with H the column height */
    double cfl;          /* dt = cfl * h / c0 */
    double delta, alpha; /* delta-SPH coefficient, artificial viscosity */
    double gamma;        /* Tait exponent (7) */
    double grid_pad;     /* padding of the cell grid around the tank (0.1 m in the reference) */
    double cell_over_kh; /* cell edge / (2h); 1 reproduces the reference grid */
} sph_dambreak_cfg;

/* ---------------------------------------------------------------- runtime */
/* sizeof of the ABI structs as compiled into the library (binding check). */
int sph_abi_sizes(int* params_size, int* cfg_size, int* tuning_size);
const char* sph_last_error(void);
int sph_device_count(int* count);
int sph_set_device(int device);
int sph_set_stream(void* hip_stream);          /* stream used by the reference-surface calls (default: null stream) */
int sph_device_synchronize(void);
int sph_malloc(void** dptr, unsigned long long bytes);
int sph_free(void* dptr);
int sph_memcpy_htod(void* dst, const void* src, unsigned long long bytes);
int sph_memcpy_dtoh(void* dst, const void* src, unsigned long long bytes);
int sph_memcpy_dtod(void* dst, const void* src, unsigned long long bytes);
int sph_memset(void* dptr, int value, unsigned long long bytes);

/* ------------------------------------------------------------- parameters */
/* Replaces cudaMemcpyToSymbol(par, temps, sizeof(SimParams)) (src/sph/sph_run.cu:105). */
int sph_set_params(const sph_params* p);
int sph_get_params(sph_params* p);
/* The reference's host recipe (src/sph/sph_run.cu:11-22, 60-101): Kleefsman tank
   3.221 x 1 x 1, cs = 25, dt = 2e-4, h = 1.5 dx, delta = 0.2, afa = 0.05,
   linear EOS, 0.1 m grid padding; float/double evaluation order kept so the
   derived floats are the reference's.  dx = 0.01 in the reference. */
int sph_params_reference(sph_params* p, float dx, int numParticles, int numFluid);
/* Parameters of a dam-break case (no reference counterpart). */
int sph_params_dambreak(const sph_dambreak_cfg* cfg, sph_params* p);
/* The same parameters on a cell grid divided by div (1, 2 or 4) along every
   axis of the case (no reference counterpart; the reference's
   opt_src_v4/sph/sph_run.cu:73 also searches a grid finer than 2h): gridSize /
   div (exact), grid dimensions times div (gridydim stays 1 in 2-D), same
   origin.  The neighbour search of such a grid visits 2*div+1 cells per axis
   (see stencil reach in sph_tuning.cell_div). */
int sph_params_subdivide(const sph_params* in, int div, sph_params* out);
/* Counts of a dam-break case without generating it. */
int sph_case_dambreak_count(const sph_dambreak_cfg* cfg, int* n_total, int* n_fluid);
/* Generates the lattice case into caller-owned host arrays of capacity cap:
   pos/vel AoS float[3N], rho/p float[N], type int[N].  Walls come first,
   then fluid.  Fluid at rest with hydrostatic pressure; walls at p = 0,
   rho = rho0.  Fills *p_out and the counts. */
int sph_case_dambreak(const sph_dambreak_cfg* cfg, int cap, float* pos, float* vel, float* rho,
                      float* p, int* type, sph_params* p_out, int* n_total, int* n_fluid);

/* ------------------------------------------- reference call surface (device pointers) */
/* calcHash (src/header/mesh.cuh:20, src/lib/mesh.cu:56-70): hash = z*X*Y + y*X + x of
   floorf((p - gridmin)/gridSize) per axis, idx[i] = i.  Cell coordinates are clamped
   into the grid (the reference reads out of bounds there instead). */
int sph_calcHash(const float* pos, int* particleHash, int* particleIndex, int numParticles);
/* sortParticles (src/header/mesh.cuh:23, src/lib/mesh.cu:73-86): stable ascending sort of
   the keys carrying the indices, in place.  Keys must lie in [0, numCells). */
int sph_sortParticles(int* particleHash, int* particleIndex, int numParticles);
/* sortParticles that allocates nothing: no reference counterpart (the
   reference's thrust::sort_by_key allocates its temporary storage).  For
   callers that cannot allocate device memory -- code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc.  Same
   arguments and results as sph_sortParticles, on the sph_set_stream stream,
   without waiting for the sort to finish.  Its scratch is one file-scope
   device array per GPU (48 MB, held until the process ends) sized for
   numParticles <= SPH_NOALLOC_MAX_PARTICLES; larger sorts are refused with
   SPH_E_INVAL.  Every call on a GPU shares that GPU's array, so calls that
   overlap on one GPU (other streams, other host threads) race. */
#define SPH_NOALLOC_MAX_PARTICLES (1 << 22)
int sph_sortParticles_noalloc(int* particleHash, int* particleIndex, int numParticles);
/* reorder (opt_src_v0/header/mesh.cuh:38, opt_src_v0/lib/mesh.cu:106-177): cellStart is
   set to -1 for every cell, then cellStart/cellEnd mark the sorted range of every
   non-empty cell (cellEnd of empty cells is left untouched, as in the reference), and
   the particle arrays are gathered into sorted order.  table_l[i] = table_r[idx[i]] +
   idx[i] - i tracks original ids.  oldZone/sortedZone/table_* may be null. */
int sph_reorder(const float* oldPos, const float* oldVel, const float* oldDensity,
                const float* oldPressure, const int* oldType, const int* oldZone,
                float* sortedPos, float* sortedVel, float* sortedDensity, float* sortedPressure,
                int* sortedType, int* sortedZone, const int* particleHash,
                const int* particleIndex, int* cellStart, int* cellEnd, int numParticles,
                int* table_l, const int* table_r);
/* recover (src/header/mesh.cuh:33, src/lib/mesh.cu:181-194, recoverD :162-179; src
   variant only): scatters the
   sorted arrays back to original order, row i of the sorted arrays to row
   particleIndex[i].  Same argument order as the reference; zone and sortedZone may
   both be null.  Needs no parameters; runs on the sph_set_stream stream. */
int sph_recover(float* pos, float* vel, float* density, float* pressure, int numParticles,
                const int* particleIndex, int* particleType, int* particleZone,
                const float* sortedPos, const float* sortedVel, const float* sortedDensity,
                const float* sortedPressure, const int* sortedType, const int* sortedZone);
/* SPH_NS_simpleversion (src/header/ns.cuh:4, opt_src_v1/equation/ns.cu:5-230): kernel A
   (wall pressure extrapolation, velocity divergence dofv) then kernel B (drho/dt with
   delta-SPH diffusion, dv/dt with pressure, artificial viscosity, acoustic damper and
   gravity).  Wall pressure/density are written in place.  rhop_sum and w_sum are
   accepted and unused, as in the reference from v0 on.  By default (ns mode 0)
   the kernels are the fast path's row-staged loops run on the caller's arrays
   (packed into library-owned scratch, allocated on first use and grown with n;
   results within 1e-5 of the reference order, rates for every particle as the
   reference computes them); ns mode 1 runs kernels in the reference's FP32
   operation order, which reproduce the reference bit for bit with the linear
   EOS.  The first call after a parameter change uploads the parameters and
   synchronises the stream. */
int sph_SPH_NS_simpleversion(const float* sortedPos, const float* sortedVel,
                             float* sortedDensity, float* sortedPressure, const int* sortedType,
                             float* densitydt, float* Veldt, const int* cellStart,
                             const int* cellEnd, int numParticles, const int* particleHash,
                             int timestep, float* dofv, float* rhop_sum, float* w_sum);
/* Selects the kernels behind sph_SPH_NS_simpleversion (no counterpart in the
   reference: it has one implementation): 0 (default) the row-staged fast
   loops, 1 the reference's FP32 operation order.  Process-wide. */
int sph_set_ns_mode(int mode);
int sph_get_ns_mode(int* mode);
/* Allocates, now, the scratch that sph_sortParticles and
   sph_SPH_NS_simpleversion would otherwise allocate on their first call for
   n particles and the current parameters' cells (no counterpart in the
   reference: Thrust allocates inside sort_by_key), so that no later call of
   up to n particles allocates.  Needs sph_set_params. */
int sph_reserve_scratch(int n);
/* PC_prediction (src/header/sph.cuh:26; opt_src_v0/sph/sph_run.cu:388-435): fluid only;
   saves (x, v, rho) into the tmp arrays at the sorted slot and advances half a step. */
int sph_PC_prediction(float* sortedPos, float* sortedVel, float* sortedDensity,
                      float* sortedPressure, const float* densitydt, const float* Veldt,
                      const int* particleHash, const int* particleIndex, float* Pos_tmp,
                      float* Vel_tmp, float* density_tmp, int numParticles,
                      int* sortedType);
/* PC_correction (src/header/sph.cuh:28; opt_src_v0/sph/sph_run.cu:437-484): fluid only;
   reads the tmp arrays through particleIndex (the half-step-2 sort permutation). */
int sph_PC_correction(float* sortedPos, float* sortedVel, float* sortedDensity,
                      float* sortedPressure, const float* densitydt, const float* Veldt,
                      const int* particleHash, const int* particleIndex, const float* Pos_tmp,
                      const float* Vel_tmp, const float* density_tmp, int numParticles,
                      int* sortedType);

/* ------------------------------------------------ simulation context (fast path) */
typedef struct sph_sim sph_sim;

/* Launch-shape knobs of the fast path, filled with defaults by sph_sim_create
   (row staging with neighbour data, prefetch, unroll 4; 3-D: runs of 32
   tiles per XCD; 2-D: one run of tiles per XCD) and chosen per workload
   with sph_sim_set_tuning.  None changes the results. */
typedef struct sph_tuning {
    int nbr_block;       /* threads per block of the neighbour-loop kernels: 64, 128, 192, 256 */
    int nbr_unroll;      /* candidate loads issued together: 1, 2, 4 or 8 */
    int xcd_remap;       /* 0: block b takes target tile b (XCD b % 8);
                            1: each XCD gets one contiguous run of tiles;
                            G (a power of two, 2..4096): runs of G consecutive
                            tiles are dealt to the 8 XCDs in turn */
    int flags;           /* bit 0: per-kernel HIP event timing in sph_sim_step */
    int nbr_lds;         /* 1: one-wave blocks stage the wave's candidate rows in LDS;
                            2: the same one z-plane (three rows) at a time;
                            3: one row at a time together with the neighbour
                            data the pair bodies read (no global loads in the
                            loop); 2 and 3 run the plain loop or mask groups
                            (nbr_compact 1 is ignored); nbr_block is then
                            ignored;
                            4: fine rows -- any stencil reach (cell_div 1 or 2):
                            the wave's targets are taken one cell row at a time,
                            the cell table of each stencil row is read once per
                            wave and each lane's candidate range is trimmed to
                            the cells its support reaches; candidates and their
                            neighbour data staged in LDS (nbr_compact, nbr_block
                            and nbr_list ignored) */
    int nbr_compact;     /* 1: scan candidates into per-lane LDS lists, then run the
                            pair body densely over the accepted ones;
                            2: test groups of 4 x nbr_unroll candidates into a bit
                            mask, then run the pair body over the set bits (with
                            nbr_lds 0, 2 or 3; nbr_lds 1 and nbr_prefetch ignored) */
    int nbr_prefetch;    /* 1: load a candidate's velocity/dofv together with its
                            position instead of after the cut-off test */
    int nbr_list;        /* 1: one scan records each target's neighbours, kernels A
                            and B iterate the lists (block/unroll/remap apply) */
    int cell_div;        /* 1: the neighbour search uses the instances' own cell grid;
                            2: a grid with half the cell edge (sph_params_subdivide),
                            2 * 2 + 1 cells per axis, about half the candidates of a
                            2h grid in 3-D.  Keys and sorted order are those of the
                            grid in use; the pairs are the same, visited row by row
                            (z, then y), ascending sorted index within a row, as on
                            any grid, so results match the oracle run on that grid.
                            Needs nbr_lds 4 unless 1. */
} sph_tuning;

/* Creates a context on the current device from host arrays in the layout of
   sph_case_dambreak (AoS float[3N] pos/vel).  Positions are widened to FP64. */
int sph_sim_create(const sph_params* p, int n, const float* pos, const float* vel,
                   const float* rho, const float* prs, const int* type, sph_sim** out);
/* Batched mode (no reference counterpart): K independent instances of equal
   dim -- e.g. dam breaks of different sizes, resolutions and time steps --
   stepped together, one launch per phase for all of them and one hipGraph per
   step.  params[k] and counts[k] describe instance k; the host arrays hold
   the instances back to back.  Every instance's results are bit-identical to
   running it alone in its own context.  Types must lie in 0..255. */
int sph_sim_create_batch(const sph_params* params, int K, const int* counts, const float* pos,
                         const float* vel, const float* rho, const float* prs, const int* type,
                         sph_sim** out);
int sph_sim_num_instances(sph_sim* s, int* K);
/* Parameters of instance k (numParticles holds the instance's particle count). */
int sph_sim_instance_params(sph_sim* s, int k, sph_params* p);
/* Per-instance surge front and column height (arrays of K doubles). */
int sph_sim_probe_batch(sph_sim* s, double band, double* front, double* height);
/* Times the neighbour kernels of the current state over the candidate launch
   shapes (block size, unroll, XCD remap), `reps` evaluations each, and keeps
   the fastest; the state is not advanced.  *best receives the choice and
   *best_us its kernel A + B time in microseconds (either may be null).
   Disabled in this build: returns SPH_E_STATE (the search faulted the GPU
   and the cause is not resolved); a context's default shape is the search's
   pick for the dam breaks, and sph_sim_set_tuning pins any other. */
int sph_sim_autotune(sph_sim* s, int reps, sph_tuning* best, double* best_us);
/* Device-to-device state exchange in ORIGINAL particle order (device
   pointers, reference layouts: pos/vel AoS float[3N], rho/p float[N],
   type int[N]; any of the outputs may be null).  set_state replaces the whole
   state (positions are widened to FP64, the sorted order restarts from the
   original order) and keeps any captured graph valid. */
int sph_sim_get_state_device(sph_sim* s, float* pos, float* vel, float* rho, float* p, int* type);
int sph_sim_set_state_device(sph_sim* s, const float* pos, const float* vel, const float* rho,
                             const float* p, const int* type);
int sph_sim_destroy(sph_sim* s);
int sph_sim_get_tuning(sph_sim* s, sph_tuning* t);
int sph_sim_set_tuning(sph_sim* s, const sph_tuning* t);
/* Runs nsteps full predictor-corrector steps (eager launches). */
int sph_sim_step(sph_sim* s, int nsteps);
/* Captures one full step into a hipGraph (replaces any previous capture). */
int sph_sim_capture(sph_sim* s);
/* Replays the captured step nsteps times. */
int sph_sim_replay(sph_sim* s, int nsteps);
int sph_sim_synchronize(sph_sim* s);
/* Copies the state back in ORIGINAL particle order: pos/vel AoS float[3N]
   (positions rounded from FP64), rho/p float[N], type int[N].  Any may be null. */
int sph_sim_download(sph_sim* s, float* pos, float* vel, float* rho, float* p, int* type);
/* FP64 positions in original order, AoS double[3N]. */
int sph_sim_download_pos64(sph_sim* s, double* pos);
/* Surge front (max x over in-domain fluid) and column height (max z over fluid
   with x <= band) computed on the device; values are written to host doubles
   (any sign; 0 when no particle qualifies). */
int sph_sim_probe(sph_sim* s, double band, double* front, double* height);
/* Front and height history without host waits: `intervals` times, k steps
   (replays of the captured step, or eager steps when none is captured)
   followed by the probe on the device; the host waits once at the end.
   front/height receive [intervals][K] doubles (K = instances).  No
   counterpart in the reference (its output step probes one frame at a time,
   src/io/io.cu:236-279); used to time a plain run while recording its front. */
int sph_sim_probe_history(sph_sim* s, int k, int intervals, double band, double* front, double* height);
/* Measuring probes of the reference's output step (src/io/io.cu:236-279),
   on the device.  Height gauges: for gauge g the largest z of a fluid
   particle of instance 0 with |x - gx[g]| < halfw (and |y - gy[g]| < halfw
   in 3-D) and z < zcap[g], 0 if there is none (the reference uses bands of
   half width 1.5 dx at x = 2.228 and 0.582 m, y = 0.5 m, the first capped at
   z < 0.4 m).  Point probes: the pressure of the particles with the given
   original indices (the reference's measuring nodes).  At most
   SPH_MAX_PROBES of each per call. */
#define SPH_MAX_PROBES 64
int sph_sim_gauges(sph_sim* s, int ngauges, const float* gx, const float* gy, const float* zcap, float halfw,
                   float* heights);
int sph_sim_pressure_at(sph_sim* s, int nids, const int* ids, float* p);
/* Average device time of each kernel launched by the last timed sph_sim_step /
   sph_sim_replay when timing is enabled (flags bit 0 of sph_tuning); names are
   written as a ';'-separated list.  Microseconds per launch. */
int sph_sim_kernel_times(sph_sim* s, char* names, int names_cap, double* us, int* count,
                         int cap);
/* Number of steps completed by this context. */
int sph_sim_steps_done(sph_sim* s, long long* steps);
/* Evaluates one half-step's neighbour phase on the current state without
   integrating: hash, sort, cell ranges, gather, kernel A, kernel B (the
   reference's calcHash..SPH_NS_simpleversion sequence).  Host outputs, any may
   be null: drho[N], acc AoS[3N], dofv[N], p/rho[N] after the wall update
   (original particle order); keys/perm[N] in sorted order, where perm holds
   the slot of the current state each sorted entry came from.  Rates of
   non-fluid particles are written as 0 (the integrator never reads them). */
int sph_sim_eval(sph_sim* s, float* drho, float* acc, float* dofv, float* p, float* rho, int* keys,
                 int* perm);
/* sph_sim_eval into DEVICE buffers, asynchronously on the context's stream
   (synchronize before reading them from another stream): drho[N] and acc AoS
   [3N] (kernel B's rates, gravity included) in original particle order; either
   may be null.  The state is not advanced. */
int sph_sim_eval_device(sph_sim* s, float* drho, float* acc);
/* Neighbour-pair statistics of the current state (for roofline accounting and
   tuning): candidate pairs visited and pairs inside the support, summed over
   all targets (kernel A) and over fluid targets (kernel B). */
int sph_sim_pair_stats(sph_sim* s, long long* candA, long long* hitsA, long long* candB,
                       long long* hitsB);

/* --------------------------------------------- files: initial state, frames, checkpoints
   Host-only readers and writers of the reference's on-disk formats. */
/* Particle count of the initial-state directory `dir`: the number of values
   in dir/rhop.dat. */
int sph_count_dat(const char* dir, int* n);
/* The reference's initial-state loader ini_fluid (src/io/io.cu:6-111): reads
   dir/h_pos.dat, h_vel.dat, h_ac.dat (3n values each), rhop.dat, p.dat (n
   each), particle_type.dat and particle_zone.dat (n integers each), all
   whitespace-separated text.  acc and zone may be null (files then not
   read).  A missing file or fewer values than asked for is SPH_E_INVAL with
   the file named (the reference exits the process). */
int sph_read_dat(const char* dir, int n, float* pos, float* vel, float* acc, float* rho, float* p, int* type,
                 int* zone);
/* Writes the same seven files (no reference counterpart; for preparing
   inputs).  Values are printed with 9 significant digits, so sph_read_dat
   reads back the same floats.  acc and zone may be null (written as 0). */
int sph_write_dat(const char* dir, int n, const float* pos, const float* vel, const float* acc, const float* rho,
                  const float* p, const int* type, const int* zone);
/* One legacy-VTK frame of the reference's output_fluid_file: an unstructured
   grid of n points with point data "p" and "velocity".  binary = 0: ASCII,
   values printed as "%.6e" (src/io/io.cu:164-196, byte-identical for the same
   arrays); binary = 1: big-endian float32 (opt_src_v3/io/io.cu:121-200, the
   same bytes).  rho and type may be null; given, they add SCALARS "rho" and
   "type" sections before "velocity". */
int sph_write_vtk(const char* path, int n, const float* pos, const float* vel, const float* p, const float* rho,
                  const int* type, int binary);
/* Checkpoint / restart (no reference counterpart).  sph_sim_checkpoint
   writes the context's complete state -- instance parameters, step count and
   the particle arrays in their current sorted order, positions in FP64 -- so
   that a context made by sph_sim_restore continues bit-identically to the
   one that was saved.  The file carries a 64-bit checksum; a truncated or
   altered file is rejected with SPH_E_INVAL.  Tuning is not saved (every
   tuning gives the same results). */
int sph_sim_checkpoint(sph_sim* s, const char* path);
int sph_sim_restore(const char* path, sph_sim** out);
/* Asynchronous frame output (the reference's writer thread,
   opt_src_v3/io/io.cu:98-119 with opt_src_v3/sph/sph_run.cu:268-300, made
   race-free): sph_sim_output_open gives the context two pinned host buffers,
   a copy stream and a writer thread.  sph_sim_output_frame gathers the
   current state into device staging (on the step stream, so steps that follow
   cannot overwrite it first), copies it to a free host buffer on the copy
   stream and queues dir/fluid_t=<frame>.vtk for the writer thread, then
   returns: later steps overlap the copy and the file write.  It blocks only
   while both buffers are still being written.  fields: bit 0 adds "rho",
   bit 1 adds "type".  sph_sim_output_flush waits for every queued frame and
   returns the first write error; sph_sim_output_close flushes and frees
   (sph_sim_destroy does too). */
int sph_sim_output_open(sph_sim* s, const char* dir, int binary, int fields);
int sph_sim_output_frame(sph_sim* s, int frame);
int sph_sim_output_flush(sph_sim* s, long long* frames_written);
int sph_sim_output_close(sph_sim* s);

/* --------------------------------------------- learned surrogate (no reference counterpart)
   A message-passing particle network (encoder, `layers` processor blocks of
   edge and node MLPs with residual updates, decoder; every MLP has two hidden
   layers of width 128) run and trained with BF16 MFMA kernels that accumulate
   in FP32.  Nodes are the particles, edges the pairs within the SPH support
   radius 2h.  Node inputs: the last `history` velocities (x, z), distances to
   the left, right and bottom walls (clipped at 2h), the fluid flag and the
   density deviation; outputs: the velocity and density change over k SPH
   steps, normalised by dv_scale and drho_scale.  2-D (x-z plane) only. */
typedef struct sph_gnn sph_gnn;

typedef struct sph_gnn_cfg {
    int layers;          /* message-passing blocks (10) */
    int history;         /* velocity history length (5) */
    int max_nodes;       /* capacity */
    int max_edges;       /* capacity */
    int train;           /* 1: keep per-layer activations for backward */
    unsigned seed;       /* weight initialisation */
    float v_scale;       /* velocity normalisation */
    float dv_scale;      /* output acceleration normalisation (mean dv/dt over k steps) */
    float drho_scale;    /* output density-rate normalisation (mean drho/dt over k steps) */
    float rho_scale;     /* input density deviation scale ((rho/rho0 - 1) * rho_scale) */
    float out_mean[3];   /* rate = normalised prediction * scale + out_mean (dvx/dt, dvz/dt, drho/dt);
                            the change over k steps is rate * k * dt, so one network serves
                            cases of different resolution (dt) */
    float out_clip;      /* rollouts clamp normalised predictions to [-out_clip, out_clip]
                            (the range the training targets were clipped to); <= 0: no clamp */
} sph_gnn_cfg;

typedef struct sph_rollout_stats {
    int accepted;        /* intervals advanced by the surrogate */
    int fallbacks;       /* surrogate intervals rejected and redone with k plain steps */
    long long plain_steps; /* plain SPH steps taken (history warm-up and fallbacks) */
    double max_dev;      /* largest per-particle density deviation the judge measured (network steps) */
    double mean_dev;     /* mean over judged network steps of the mean per-particle deviation */
    double seconds;      /* wall-clock of the rollout */
    int rejected_penetration; /* network steps rejected for wall penetration (a fluid particle
                                 predicted outside the escape box, where the solver would park it) */
    double max_penetration;   /* deepest predicted position beyond the escape box seen by the judge (m) */
    int rejected_density;     /* network steps rejected for a density deviation above dev_tol */
    int rejected_graph;       /* intervals left to the plain step because the state's graph had more
                                 than max_edges edges */
    int rejected_velocity;    /* network steps rejected for a velocity residual above acc_tol */
    double max_band_dev;      /* largest front-band velocity residual of network steps (units of g T) */
    double max_bias_dev;      /* largest whole-flow velocity residual of network steps (units of g T) */
    double plain_max_dev;     /* the judge's readings of the plain intervals (warm-up and fallbacks): */
    double plain_max_band_dev; /*   density deviation, front-band and whole-flow velocity residuals, */
    double plain_max_bias_dev; /*   which a threshold must clear so that plain physics would pass */
    double max_drift;         /* largest accumulated front-band velocity drift of accepted steps (g T) */
    int skipped;              /* intervals given to the plain step without a network attempt
                                 (the back-off after rejections; not judged) */
} sph_rollout_stats;

/* Advances the single-instance 2-D context `sim` by `intervals` intervals of
   T = k plain-step lengths, each by one surrogate step judged by the plain
   method.  `judge` is a second context of the same case: the plain rates
   (kernel B's acceleration and continuity rate) are evaluated in it at the
   interval's start state S and at the prediction S'.  Over the fluid
   particles inside the escape box the prediction must satisfy
   - density: |rho' - max(rho0, rho + T (drho(S) + drho(S')) / 2)| / rho0 <=
     dev_tol for every particle (the continuity integral, trapezoid rule);
   - velocity: with r = v' - v - T (a(S) + a(S')) / 2, the mean of r over the
     front band (particles with x >= front - front_band, front = max x of
     fluid at the start) and the mean over all fluid, each component, are at
     most acc_tol g T in magnitude (a lagging or leading surge front, or a bias
     of the whole flow), and the front band's signed mean residual (x and z),
     summed over the accepted network steps including this one, stays within
     drift_tol g T (a small lag of the front repeated over many steps adds up;
     plain intervals add nothing);
   - no fluid particle is predicted outside the escape box (the tank faces;
     the solver would park it).
   Otherwise the interval is redone with k plain steps (bit-identical to the
   plain run), and with max_skip > 0 the next 1, 2, 4, ... (at most max_skip)
   intervals after a rejection go to the plain step without a network
   attempt (the wait doubles with every rejection and clears with an
   acceptance); max_skip 0 attempts the network at every interval.  Particles the solver has already parked are left where they
   are and out of the judgement.  The first `history` intervals are plain (they
   fill the velocity history).  front/height (may be null) receive the probes
   (band: column-height band) after every interval.  progress (may be null)
   is called every 1000 intervals with the intervals done and accepted. */
typedef void (*sph_progress_fn)(int done, int accepted, void* user);
int sph_gnn_rollout(sph_gnn* g, sph_sim* sim, sph_sim* judge, int k, int intervals, float dev_tol, float acc_tol,
                    float drift_tol, double front_band, double band, int max_skip, double* front, double* height,
                    sph_rollout_stats* stats, sph_progress_fn progress, void* user);

/* The rollout's judge on one given interval of k steps (host arrays, original
   order, 2-D single-instance `judge` context of the case): start state
   pos/vel/rho/p/type, end state pos2/vel2/rho2/p2.  out[7]: largest and mean
   density deviation, front-band and whole-flow velocity residuals (units of
   g T), fluid particles outside the escape box, deepest position beyond it,
   front-band particle count.  The judge context's state is replaced. */
int sph_gnn_judge(sph_sim* judge, int k, double front_band, const float* pos, const float* vel, const float* rho,
                  const float* p, const int* type, const float* pos2, const float* vel2, const float* rho2,
                  const float* p2, double* out);

int sph_gnn_create(const sph_gnn_cfg* cfg, sph_gnn** out);
int sph_gnn_destroy(sph_gnn* g);
int sph_gnn_num_params(sph_gnn* g, long long* count);
/* FP32 master parameters (flat, in the layout of sph_gnn_param_layout) */
int sph_gnn_get_params(sph_gnn* g, float* host);
int sph_gnn_set_params(sph_gnn* g, const float* host);
int sph_gnn_get_grads(sph_gnn* g, float* host);
/* Layout of the flat parameter vector: for every linear layer (in the order
   node encoder, edge encoder, [edge MLP, node MLP] x layers, decoder; three
   layers per MLP) the five ints {weight offset, N, K, bias offset, 0};
   W is N x K row-major, b has N entries.  *nlayers receives the count. */
int sph_gnn_param_layout(sph_gnn* g, int* layout5, int cap, int* nlayers);
/* Loads one particle state (host arrays, original order): pos AoS[3n],
   hist[history][n][3] newest first, rho[n], type[n]; target[n][3] (normalised
   output) may be null.  Builds the radius graph on the device. */
int sph_gnn_load_state(sph_gnn* g, const sph_params* p, int n, const float* pos, const float* hist,
                       const float* rho, const int* type, const float* target);
/* Graph of the loaded state: edge count and, if the pointers are not null,
   CSR offsets[n+1], senders[e] and receivers[e]. */
int sph_gnn_graph(sph_gnn* g, int* edges, int* offsets, int* senders, int* receivers);
/* Node and edge input features (128 wide, FP32) of the loaded state. */
int sph_gnn_features(sph_gnn* g, float* node_feat, float* edge_feat);
/* Activations of the last forward pass (needs train = 1), for checking the
   backward pass against a CPU restatement evaluated at the same activations.
   kind 0/1: hidden activations h1/h2 of MLP call `index` (BF16 as uint16,
   rows x 128; calls: 0 node encoder, 1 edge encoder, 2 + 2l edge update and
   3 + 2l node update of block l, 2 + 2L decoder; rows = edges for the edge
   encoder and edge updates, nodes otherwise); 2: node latents V_index
   (n x 128 FP32, index 0..L); 3: edge latents E_index (edges x 128);
   4: aggregated messages A_index (n x 128, index 0..L-1); 5: decoder output
   (n x 128 FP32, columns 3.. are zero).  count must equal the element count. */
int sph_gnn_get_tensor(sph_gnn* g, int kind, int index, void* host, long long count);
/* Forward pass; out[n][3] (normalised outputs) may be null; *loss gets the
   mean squared error over fluid nodes when a target was loaded. */
int sph_gnn_forward(sph_gnn* g, float* out, double* loss);
/* Backward pass of the last forward (needs train = 1); accumulates into the gradients. */
int sph_gnn_backward(sph_gnn* g);
int sph_gnn_zero_grad(sph_gnn* g);
/* Adam step (beta 0.9 / 0.999, eps 1e-8) with global-norm clipping at clip (<= 0: none). */
int sph_gnn_adam(sph_gnn* g, float lr, float clip);
/* Forward MLP calls as one fused kernel each (fused = 1, 2 or 4: hidden
   activations stay in LDS, operand loads run `fused` k-steps ahead of the
   MFMAs; the default with train = 1) or as three GEMM launches (fused = 0);
   these settings give bit-identical results.  fused = 8 (train = 0 only):
   the edge MLPs apply their first layer's receiver and sender blocks to the
   nodes first and run as one persistent kernel with the weights resident in
   LDS, which also aggregates the messages -- the same products summed in
   another order, so outputs agree with the other settings to FP32 rounding
   (and the BF16 rounding of hidden activations that it can flip).  fused = 9
   (train = 0 only, the default there): as 8 with the edge latents stored as
   BF16 between blocks (halving their HBM traffic; outputs agree with 8 to
   about 1e-2). */
int sph_gnn_set_fused(sph_gnn* g, int fused);
/* One MFMA tile D(16x16) = A(16x32) B(32x16) in BF16 (host arrays as raw
   uint16 BF16 bits), for checking the operand layout. */
int sph_gnn_mfma_probe(const unsigned short* A, const unsigned short* B, float* D);

#ifdef __cplusplus
}
#endif
#endif /* SPH_CAPI_H */
