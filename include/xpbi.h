/*
 * xpbi.h -- C ABI of the MI355X XPBI substep engine (libxpbi.so).
 *
 * The reference drives its substep from JavaScript through WebGPU:
 *   class Simulation            web/js/sim.js:9-273
 *   constructor(device, scene)  web/js/sim.js:10-30 (grid set-up)
 *   createBuffers()             web/js/sim.js:32-82 (state upload)
 *   writeMaterials(materials)   web/js/sim.js:84-99 (64-B records)
 *   writeParams(p)              web/js/sim.js:159-186 (256-B record)
 *   encodeSubstep(enc, I, w)    web/js/sim.js:220-272 (one substep)
 *   profileSubstep(I, w)        web/js/sim.js:199-215 (per-pass timing)
 *   readPos / readVel hooks     web/js/main.js:291-312
 * Each entry point below replaces one of those; the Node N-API addon and the
 * Python ctypes binding wrap them 1:1.  All functions return XPBI_OK (0) or a
 * negative XPBI_ERR_* code; xpbi_last_error() gives the message.
 *
 * Threading: a sim owns one HIP stream; calls on one sim are not reentrant.
 * xpbi_substep / xpbi_step enqueue asynchronously; xpbi_read synchronises.
 * After every substep the current velocity is the "velA" of the reference
 * (web/js/sim.js:217, :268-271).
 */
#ifndef XPBI_H
#define XPBI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XPBI_OK 0
#define XPBI_ERR_ARG (-1)     /* bad argument (null, size mismatch, range) */
#define XPBI_ERR_HIP (-2)     /* HIP runtime error */
#define XPBI_ERR_STATE (-3)   /* call out of order (e.g. substep before params) */
#define XPBI_ERR_NOMEM (-4)   /* device allocation failed */

/* Material kinds, web/js/sim.js:5-7 and web/js/shaders/common.js:13-19 */
enum {
  XPBI_MAT_ELASTIC = 0,
  XPBI_MAT_DP = 1,
  XPBI_MAT_NACC = 2,
  XPBI_MAT_VM = 3,
  XPBI_MAT_HB = 4,
  XPBI_MAT_SNOW = 5,
  XPBI_MAT_WATER = 6
};

/* 64-byte material record: struct Material, web/js/shaders/common.js:21-33,
 * packed as web/js/sim.js:84-99 does (mu, lambda from E, nu on the host). */
typedef struct xpbi_material {
  float mu, lam, density;
  uint32_t kind;
  float p0, p1, p2, p3; /* DP(alpha,coh) NACC(alpha0,beta,xi,M) VM(sigY) HB(sigY,h,eta) SNOW(thC,thS,xi) */
  float color[4];
  float boundFriction;
  float mass; /* density * V0 */
  float pad0, pad1;
} xpbi_material;

/* 256-byte parameter record: struct Params, web/js/shaders/common.js:35-60,
 * word offsets as written by web/js/sim.js:159-186. */
typedef struct xpbi_params {
  float gridOrigin[3];
  float cellSize;
  int32_t gridDims[3];
  uint32_t numCells;
  float domainMin[3];
  float dt;
  float domainMax[3];
  float h;
  float gravity[3];
  float support;
  float colliderPos[3];
  float colliderR;
  float colliderVel[3];
  float colliderActive;
  uint32_t N;
  float V0, xsphC, omega, pbfEps, sCorrK, time, contactOmega, kDamp;
  float pad[27];
} xpbi_params;

/* Scene description: the scene object of web/js/scenes.js:94-176 as consumed
 * by the Simulation constructor and createBuffers (web/js/sim.js:10-82). */
typedef struct xpbi_scene_desc {
  uint32_t N;
  const float* positions;  /* 3N */
  const float* tags;       /* N: pos.w = matIdx + colour fraction < 0.49 (sim.js:63); NULL -> matIds */
  const uint32_t* matIds;  /* N, used when tags == NULL */
  const float* states;     /* N plastic state (aux.z), NULL -> 0 */
  const float* velocities; /* 3N, NULL -> 0 */
  double h;                /* particle spacing / kernel radius; support = 2h */
  double domainMin[3];
  double domainMax[3];
  int device; /* HIP device ordinal */
} xpbi_scene_desc;

/* Fields for xpbi_read / xpbi_write, in the reference's buffer layouts
 * (web/js/sim.js:32-50, web/js/shaders/simkernels.js:18-57). */
enum {
  XPBI_FIELD_POS = 0,       /* 16N: vec4 (x, y, z, tag) */
  XPBI_FIELD_VEL = 1,       /* 16N: vec4 (v, w = V^n after scatter) -- velA */
  XPBI_FIELD_F = 2,         /* 48N: 3 columns, each a vec4 with pad 0 */
  XPBI_FIELD_L = 3,         /* 48N: same layout, read only */
  XPBI_FIELD_AUX = 4,       /* 16N: (V^n, lambda, plastic state, pbf lambda) */
  XPBI_FIELD_SP = 5,        /* 48N: columns of dlam*G*L, selfDv in the w lanes; read only */
  XPBI_FIELD_CELLSTART = 6, /* 4*numCells: exclusive start offsets (cellTab after scatter); read only */
  XPBI_FIELD_PERM = 7,      /* 4N: sortedIdx of the last sort; read only */
  XPBI_FIELD_ID = 8,        /* 4N: persistent particle id, permuted with the particles */
  XPBI_FIELD_MATS = 9,      /* 512 per scene: the 8 64-B material records each scene's kernels
                               read (matsBuf, web/js/sim.js:49; MaterialR, render.js:16-21); read only */
  XPBI_FIELD_COUNT = 10
};

/* Stages for xpbi_substep_partial (staged parity checks); XPBI_STAGE_FULL is a
 * whole substep.  Order follows web/js/sim.js:246-266. */
enum {
  XPBI_STAGE_SORT = 1,
  XPBI_STAGE_REORDER = 2,
  XPBI_STAGE_COMPUTE_L = 3,
  XPBI_STAGE_SOLVE = 4,
  XPBI_STAGE_XSPH = 5,
  XPBI_STAGE_UPDATE_F = 6,
  XPBI_STAGE_FULL = 7
};

typedef struct xpbi_sim xpbi_sim;

/* new Simulation(device, scene) + createBuffers(): web/js/sim.js:10-82.
 * Grid: cs = 2h, origin = domainMin - 2cs, dims = ceil((max-min)/cs) + 4
 * (sim.js:20-26), computed in double like the JavaScript host. F = I,
 * aux = (h^3, 0, state, 0), id = 0..N-1. Materials must be written next. */
int xpbi_create(const xpbi_scene_desc* desc, xpbi_sim** out);

/* Frees every device buffer (the reference has no destroy; sim.js:9). */
void xpbi_destroy(xpbi_sim* sim);

/* Message of the last failed call on sim (sim == NULL: last create failure). */
const char* xpbi_last_error(const xpbi_sim* sim);

/* Grid as computed at create (the JavaScript host packs it into Params the
 * way web/js/sim.js:174-176 does); scene 0 of a batch. */
int xpbi_grid_info(const xpbi_sim* sim, float origin[3], float* cell_size, int32_t dims[3], uint32_t* num_cells,
                   uint32_t* n_particles);

/* writeMaterials(materials): web/js/sim.js:84-99; n <= 8 records (scene 0). */
int xpbi_write_materials(xpbi_sim* sim, const xpbi_material* mats, int n);

/* writeParams(p): web/js/sim.js:159-186.  The grid words and N must equal the
 * sim's own (xpbi_grid_info); support must not exceed cellSize (scene 0). */
int xpbi_write_params(xpbi_sim* sim, const xpbi_params* p);

/* ---- Batches (no reference counterpart: the reference runs one Simulation per
 * scene, web/js/sim.js:9-30).  `count` independent scenes of any sizes share one
 * simulation: their particles are concatenated in scene order (scene q holds
 * particles first_particle .. first_particle + n_particles - 1 of every field),
 * their cell grids are stacked in one cell table (CELLSTART has the sum of the
 * scenes' cells; PERM indexes the concatenated arrays), and every kernel of a
 * substep is one launch for all scenes.  Scenes never interact.  Each scene
 * has its own materials and params, written with the scene variants below; the
 * Jacobi residual of xpbi_set_tolerance is taken over the whole batch.  A single
 * scene (xpbi_create) is a batch of one. */
int xpbi_create_batch(const xpbi_scene_desc* descs, int count, xpbi_sim** out);
int xpbi_scene_count(const xpbi_sim* sim, int* count);
int xpbi_scene_info(const xpbi_sim* sim, int scene, float origin[3], float* cell_size, int32_t dims[3],
                    uint32_t* num_cells, uint32_t* n_particles, uint32_t* first_particle, uint32_t* first_cell);
int xpbi_write_scene_materials(xpbi_sim* sim, int scene, const xpbi_material* mats, int n);
int xpbi_write_scene_params(xpbi_sim* sim, int scene, const xpbi_params* p);

/* Launch shape of the particle kernels (no reference counterpart).  The
 * particles are cut into runs of 256 that never cross a scene (the work
 * entries); blocks = 0 launches one block per entry, blocks > 0 a persistent
 * grid of that many blocks (rounded up to a multiple of 8, at most one per
 * entry) in which each XCD walks a contiguous eighth of the entries. */
int xpbi_set_launch(xpbi_sim* sim, int blocks);

/* Slots of the support-neighbour list (no reference counterpart: the reference
 * walks the 9 cell rows in every sweep, web/js/shaders/simkernels.js:59-74, and
 * this library records the neighbours once per substep instead).  Default 256, at
 * 2 bytes per slot per particle; 0..1024, rounded up to a multiple of 16.  A particle with more neighbours walks
 * its cell rows, and slots = 0 walks rows everywhere and frees the list: the
 * results are the same bit for bit, only the speed and the memory differ. */
int xpbi_set_neighbour_slots(xpbi_sim* sim, int slots);
int xpbi_get_launch(const xpbi_sim* sim, int* blocks, int* work_entries);

/* substeps x substep like xpbi_step, replayed from a HIP graph of two captured
 * substeps (captured on first use and again whenever the iteration count, water
 * flag, tolerance, warm-start mode, launch shape or residual monitor changes;
 * host polling of xpbi_set_tolerance is off inside the graph).  An odd
 * remainder runs as an ordinary substep.  Results equal xpbi_step's. */
int xpbi_graph_step(xpbi_sim* sim, int substeps, int iterations, int has_water);

/* Autotuner: times graph-replayed substeps under each launch shape (one block
 * per entry; persistent grids of 1, 2 and 4 blocks per compute unit), keeps the
 * fastest, and restores the particle state it started from.  best_blocks gets
 * the chosen xpbi_set_launch value, best_ms its time per substep. */
int xpbi_autotune(xpbi_sim* sim, int iterations, int has_water, int substeps, int* best_blocks, float* best_ms);

/* encodeSubstep(enc, iterations, hasWater): web/js/sim.js:220-272. */
int xpbi_substep(xpbi_sim* sim, int iterations, int has_water);

/* The frame loop's substep chunk (web/js/main.js:406-413): substeps x substep. */
int xpbi_step(xpbi_sim* sim, int substeps, int iterations, int has_water);

/* A substep that stops after `stage` (XPBI_STAGE_*), for staged parity. */
int xpbi_substep_partial(xpbi_sim* sim, int iterations, int has_water, int stage);

/* readPos/readVel (web/js/main.js:291-312) generalised to every buffer;
 * bytes must equal the field's size.  Synchronises the sim's stream. */
int xpbi_read(xpbi_sim* sim, int field, void* dst, size_t bytes);

/* Overwrite a state field (POS, VEL, F, AUX, ID) between substeps. */
int xpbi_write(xpbi_sim* sim, int field, const void* src, size_t bytes);

/* Wait for all enqueued work of the sim. */
int xpbi_sync(xpbi_sim* sim);

/* The device buffers a renderer binds (Renderer.setScene, web/js/main.js:189-199:
 * posBuf, matsBuf and velA), for zero-copy interop.  pos and vel are N float4
 * (XPBI_FIELD_POS / _VEL layouts), mats is 8 material records per scene
 * (XPBI_FIELD_MATS), stream the hipStream_t the sim enqueues on: a consumer
 * orders its reads after the sim's work on it (or after xpbi_sync).  pos and
 * vel move between two buffers as the particles are re-sorted, so the view is
 * valid until the next call that steps, writes or destroys the sim. */
typedef struct xpbi_render_view {
  const void* pos;
  const void* vel;
  const void* mats;
  void* stream;
  uint32_t N;
  uint32_t scenes;
  int device;
} xpbi_render_view;
int xpbi_get_render_view(xpbi_sim* sim, xpbi_render_view* out);

/* profileSubstep(iterations, hasWater): web/js/sim.js:199-215.  Runs one
 * substep with a HIP event pair around every kernel launch; writes up to cap
 * per-launch milliseconds to ms and their count to *count; labels via
 * xpbi_profile_label.  Repeated launches (solve iterations) are listed each. */
int xpbi_profile_substep(xpbi_sim* sim, int iterations, int has_water, float* ms, int cap, int* count);
const char* xpbi_profile_label(const xpbi_sim* sim, int k);

/* Enqueue `substeps` substeps between two HIP events on the sim's stream and
 * return the elapsed milliseconds (inputs stay resident in HBM). */
int xpbi_time_substeps(xpbi_sim* sim, int substeps, int iterations, int has_water, float* ms);

/* As xpbi_time_substeps, with a HIP event after every kernel launch inside the
 * timed region; per distinct label (xpbi_profile_label(sim, k), k < *nlabels)
 * label_ms[k] is the summed milliseconds and label_count[k] the launches. */
int xpbi_time_substeps_detail(xpbi_sim* sim, int substeps, int iterations, int has_water, float* total_ms,
                              float* label_ms, int* label_count, int cap, int* nlabels);

/* Stream-copy ceiling: copies `bytes` (rounded down to 16) between two device
 * buffers allocated for the call with a float4 copy kernel, `reps` times, and
 * returns the milliseconds of the last copy.  No reference counterpart; it is
 * the measured bandwidth ceiling and the known-size kernel that HBM byte
 * counters are calibrated against. */
int xpbi_copy_bandwidth(xpbi_sim* sim, size_t bytes, int reps, float* ms);

/* Residual monitor (no reference counterpart: the reference runs a fixed
 * iteration count, web/js/sim.js:256).  When enabled, every solveA reduces
 * r_k = ||C + alpha~ lambda + gamma dt dC.v||_2 over solid particles; after a
 * substep, out[2k] = sum of squares at iteration k, out[2k+1] = sum C^2 at k. */
int xpbi_set_residual_monitor(xpbi_sim* sim, int enabled);
int xpbi_read_residuals(xpbi_sim* sim, double* out, int cap, int* count);

/* Residual-driven Jacobi loop (replaces the fixed count of web/js/sim.js:256;
 * no reference counterpart).  With tol > 0, the `iterations` argument of
 * xpbi_substep / xpbi_step is a cap: iteration k computes
 * r_k = ||C + alpha~ lambda + gamma dt dC.v||_2 / ||C at iteration 0||_2 over solid
 * particles (the numerator of Delta-lambda / omega, simkernels.js:305) and is the
 * last one applied once r_k <= tol; the test runs on the device, so a substep
 * needs no host synchronisation.  check_every > 0 also lets the host stop
 * enqueuing iterations: every check_every iterations it reads the device's
 * report (written into pinned host memory, no stream synchronisation) and waits
 * there until the device is at most 2 iterations behind, so that launches past
 * the last iteration (which return at entry) stay few.  tol = 0 restores the
 * fixed count. */
int xpbi_set_tolerance(xpbi_sim* sim, float tol, int check_every);
/* Jacobi iterations applied by the last substep (a learned first iteration
 * counts as one) and, with tol > 0, its final residual r_k (-1 otherwise).
 * Synchronises. */
int xpbi_last_iterations(xpbi_sim* sim, int* iterations, double* residual);

/* ---- Learned warm start (no reference counterpart: the reference starts every
 * substep's multipliers at zero, simkernels.js:146, and bisects the HB flow
 * rule over its full bracket, common.js:264-265).
 *
 * A 12 -> 64 -> 64 -> 2 MLP (BF16 weights and activations, f32 accumulation;
 * canonical parameter order W1[64][12], b1[64], W2[64][64], b2[64], W3[2][64],
 * b3[2], row = output unit) maps each solid particle's features at lambda = 0 to
 * (rho, t): the warm start replaces the first Jacobi iteration's step
 * Delta-lambda = omega d0 (d0 = (-C - gamma dt dC.v) / ((1 + gamma) denom + alpha~)
 * at lambda = 0) by lambda0 = clamp(rho, 0, 2) d0 (at most twice the Newton step,
 * whatever the network outputs) and applies it like any
 * iteration (solveB, and PBF with water); plain Jacobi then continues, to the
 * tolerance of xpbi_set_tolerance measured against the same ||C|| at
 * lambda = 0 as a plain run, or for the remaining iterations of a fixed count
 * (the learned iteration counts as one).  A network that outputs rho = omega
 * reproduces plain Jacobi exactly.  t is the HB return map's predicted root
 * position in its bisection bracket.  Features (x, 12 per particle):
 * 100 log sigma of the trial F (3, descending), 100 tr log sigma and
 * 100 |dev log sigma| after the return map, log10(mu) - 5, lambda / (lambda + 2 mu),
 * kind / 6, material strength, support neighbours / 40,
 * clamp(previous lambda / d0, +-8), alpha~ / ((1 + gamma) denom + alpha~). */
#define XPBI_MLP_INPUTS 12
#define XPBI_MLP_OUTPUTS 2
#define XPBI_MLP_PARAMS 5122

/* Enable the warm start with the given network (a copy is kept), or disable it
 * with params = NULL. */
int xpbi_set_warmstart(xpbi_sim* sim, const float* params, size_t n);
/* Run one plain substep (like xpbi_substep) and return, per particle in the
 * order of the substep, its features x (rows x 12), training targets t
 * (rows x 2): lambda at the end of the solve / d0 and the HB root position of
 * the first iteration's return map (NaN where there is none), and w (rows, may
 * be NULL): the particle's weight in the residual, (-C - gamma dt dC.v)^2 at
 * lambda = 0 (0 without an active constraint).  rows = N. */
int xpbi_collect_warmstart_pairs(xpbi_sim* sim, int iterations, int has_water, float* x, float* t, float* w,
                                 size_t rows);
/* Start every substep's Jacobi from given multipliers (no counterpart in the
 * reference: the warm start's insertion point, simkernels.js:302-307, fed by the
 * caller, e.g. with the multipliers an earlier solve converged to): from the
 * next substep on, the first iteration applies Delta-lambda = lambda0[id] to
 * each solid particle whose constraint is active (id: its persistent id,
 * XPBI_FIELD_ID), then plain Jacobi follows.  n = N copies lambda0 to the
 * device; lambda0 = NULL with n = 0 turns it off.  It replaces a network given
 * to xpbi_set_warmstart, and a network given later replaces it. */
int xpbi_set_lambda0(xpbi_sim* sim, const float* lambda0, size_t n);

/* The MLP trainer: forward and backward on BF16 MFMA, Adam (beta 0.9, 0.999,
 * eps 1e-8) on f32 master weights.  Loss = mean over rows of
 * w_row sum_k (y_k - t_k)^2, skipping NaN targets (w = 1 without weights).  One trainer per device; calls on one
 * trainer must not overlap. */
typedef struct xpbi_mlp xpbi_mlp;
int xpbi_mlp_create(int device, xpbi_mlp** out);
void xpbi_mlp_destroy(xpbi_mlp* mlp);
const char* xpbi_mlp_last_error(const xpbi_mlp* mlp);
/* Set the f32 master weights (resets the Adam moments and step count). */
int xpbi_mlp_set_params(xpbi_mlp* mlp, const float* params, size_t n);
int xpbi_mlp_get_params(xpbi_mlp* mlp, float* params, size_t n);
/* Training data: x rows x 12, t rows x 2, optional sample weights w (rows, NULL:
 * all 1); copied to the device. */
int xpbi_mlp_set_data(xpbi_mlp* mlp, const float* x, const float* t, const float* w, size_t rows);
/* y (rows x 2) = the network on x (rows x 12). */
int xpbi_mlp_predict(xpbi_mlp* mlp, const float* x, size_t rows, float* y);
/* Loss and its gradient (canonical order) over data rows [first, first + count). */
int xpbi_mlp_loss_grad(xpbi_mlp* mlp, size_t first, size_t count, float* grad, float* loss);
/* One Adam step with learning rate lr on grad (canonical order), or on the
 * gradient of the last xpbi_mlp_loss_grad when grad = NULL. */
int xpbi_mlp_adam_step(xpbi_mlp* mlp, const float* grad, float lr);
/* Minibatch Adam over the data: each epoch visits the rows in a fresh order
 * (seeded); epoch_loss (epochs floats, may be NULL) gets each epoch's mean loss. */
int xpbi_mlp_train(xpbi_mlp* mlp, int epochs, int batch, float lr, uint64_t seed, float* epoch_loss);

/* ---- Allocation-free substep.  No counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  One substep of xpbi_substep
 * (web/js/sim.js:220-272; fixed iteration count, no warm start) of a single scene
 * whose state lives in the caller's device buffers, in the reference layouts of
 * xpbi_read: pos (16N: x, y, z, tag), vel (16N: velA), F (48N: three vec4 columns)
 * and aux (16N), each 16-byte aligned; the substep reads them and overwrites them
 * with its result, the particles in cell order as the reference's reorder leaves
 * them.  p is the 256-byte Params record (p->N particles, grid words as
 * xpbi_create computes them); mats holds n <= 8 64-byte records, mass filled
 * (xpbi_write_materials).  Work is enqueued on `stream` (a hipStream_t; NULL: the
 * null stream) of the current device, and the call returns without waiting.
 * Every intermediate (sorted copies, cell table, sort keys, neighbour list, L,
 * Sp) lives in a static device array of the library of
 * xpbi_noalloc_scratch_bytes() bytes, one per device, held until the process
 * ends; the call allocates nothing.  Limits: N <= XPBI_NOALLOC_MAX_PARTICLES and
 * numCells <= XPBI_NOALLOC_MAX_CELLS, else XPBI_ERR_ARG (message from
 * xpbi_last_error(NULL)).  All calls on one device share that array: calls that
 * overlap on one device (other streams, other threads) race. */
#define XPBI_NOALLOC_MAX_PARTICLES 65536u
#define XPBI_NOALLOC_MAX_CELLS 1048576u
int xpbi_substep_noalloc(const xpbi_params* p, const xpbi_material* mats, int n, void* pos, void* vel, void* F,
                         void* aux, int iterations, int has_water, void* stream);
size_t xpbi_noalloc_scratch_bytes(void);

/* ---- 2D fp64 engine: the algorithm of prototype/xpbi2d.py:123-241 ---------
 * The prototype is a script (no API): a sand column whose state is x, v, F
 * (xpbi2d.py:52-58) advanced by the main loop (xpbi2d.py:123-241) with
 * constants fixed at xpbi2d.py:29-49.  Here the constants are xpbi2d_params and
 * the loop is xpbi2d_step, which runs a call's substeps in one launch, in
 * float64. */
typedef struct xpbi2d_params {
  double h;           /* spacing = kernel radius (xpbi2d.py:29-30); support 2h, V0 = h^2 */
  double E, nu, rho;  /* StVK-Hencky moduli and density (xpbi2d.py:32-36) */
  double phi_deg;     /* Drucker-Prager friction angle (xpbi2d.py:37-38) */
  double dt;          /* substep (xpbi2d.py:39) */
  int iterations;     /* Jacobi iterations (xpbi2d.py:40) */
  double gravity[2];  /* (xpbi2d.py:41) */
  double xsph_c;      /* (xpbi2d.py:42) */
  double domain[2];   /* box [0, domain]; walls at h from each side (xpbi2d.py:49, :205-206) */
  double omega;       /* Delta-lambda multiplier; the prototype's is 1 (xpbi2d.py:178-180) */
} xpbi2d_params;
enum {
  XPBI2D_FIELD_X = 0,      /* 2N doubles: positions */
  XPBI2D_FIELD_V = 1,      /* 2N doubles: velocities */
  XPBI2D_FIELD_F = 2,      /* 4N doubles: F row-major (F00, F01, F10, F11) */
  XPBI2D_FIELD_LAMBDA = 3  /* N doubles: the last substep's lambda; read only */
};
#define XPBI2D_MAX_PARTICLES 65536u
typedef struct xpbi2d_sim xpbi2d_sim;
/* positions: 2N doubles; v = 0 and F = I initially (xpbi2d.py:57-58). */
int xpbi2d_create(uint32_t N, const double* positions, const xpbi2d_params* params, int device, xpbi2d_sim** out);
void xpbi2d_destroy(xpbi2d_sim* sim);
const char* xpbi2d_last_error(xpbi2d_sim* sim);
/* substeps iterations of xpbi2d.py:123-237; returns when they are done.
 * XPBI_ERR_STATE when a particle had more than 48 neighbours within 2h. */
int xpbi2d_step(xpbi2d_sim* sim, int substeps);
/* How a scene's substeps run: single = 0 (the default where the device supports
 * cooperative launches) a grid of 256-thread workgroups, one thread per particle,
 * separated by grid barriers; single = 1 one workgroup of 1024 threads.  Results
 * are the same bit for bit. */
int xpbi2d_set_workgroups(xpbi2d_sim* sim, int single);
/* xpbi2d_step with a HIP event pair around the launch: *ms = its duration. */
int xpbi2d_time(xpbi2d_sim* sim, int substeps, float* ms);
/* count = the field's size in doubles. */
int xpbi2d_read(xpbi2d_sim* sim, int field, double* dst, size_t count);
int xpbi2d_write(xpbi2d_sim* sim, int field, const double* src, size_t count);

#ifdef __cplusplus
}
#endif

#endif /* XPBI_H */
