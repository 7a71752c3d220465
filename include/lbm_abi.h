/*
 * C-ABI of the MI355X (gfx950) D2Q9 lattice-Boltzmann hot path.
 *
 * The reference (Chiuzzmm/taichi-lbm) has no FFI: its operator interface is
 * duck-typed Python objects whose @ti.kernel methods mutate Taichi fields.
 * Each entry point below replaces one such kernel (cited file:line relative to
 * the reference tree) or a fused sequence of them, and is bound from Python
 * with ctypes (taichi_lbm/_hip.py).
 *
 * Conventions
 *  - Every function returns 0 on success or a negative LBM_E* status; the
 *    thread-local text of the last failure is returned by lbm_last_error().
 *    Nothing is thrown or aborted across the ABI.
 *  - All buffers are device pointers owned by the caller; the library never
 *    allocates or frees them.  `stream` is a hipStream_t passed as void*
 *    (NULL = the null stream).  Calls are asynchronous unless stated.
 *  - Field layout ("plane" layout): a scalar field over the NX x NY lattice is
 *    one plane of NX*NY values, node (i, j) at offset i*NY + j (i = x slowest,
 *    the order of the reference's (NX, NY) Taichi fields).  Multi-valued fields
 *    are consecutive planes `plane_stride` elements apart:
 *      populations  plane (c*9 + k)            c = component, k = D2Q9 index
 *      macro        planes rho_0..rho_{nc-1}, total_rho, ux, uy
 *      vectors      plane (c*2 + axis)
 *  - D2Q9 numbering is the reference's (lbm.py:43-45):
 *      c = (0,0) (1,0) (0,1) (-1,0) (0,-1) (1,1) (-1,1) (-1,-1) (1,-1).
 */
#ifndef TAICHI_LBM_AMD_LBM_ABI_H
#define TAICHI_LBM_AMD_LBM_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 17

/* status codes */
#define LBM_OK 0
#define LBM_E_INVALID (-1)  /* bad argument / unsupported combination */
#define LBM_E_HIP (-2)      /* a HIP runtime call failed */

/* collision models (CollisionEngine.py:36-451) */
enum { LBM_BGK = 0, LBM_TRT = 1, LBM_MRT = 2 };

/* population storage formats */
enum {
  LBM_STORE_F32 = 0,       /* f itself, fp32 */
  LBM_STORE_F16_SHIFT = 1, /* f - w_k, IEEE half; arithmetic in fp32 */
  LBM_STORE_BF16_SHIFT = 2, /* f - w_k, bfloat16; arithmetic in fp32 */
  LBM_STORE_F64 = 3         /* f itself, fp64; arithmetic in fp64 (high-precision path) */
};

/* rounding of a step's populations to FP16 / BF16 storage (lbm_lattice.rounding) */
enum {
  LBM_ROUND_NEAREST = 0,    /* round to nearest even */
  LBM_ROUND_STOCHASTIC = 1  /* gfx950 stochastic conversions: the expected stored value is the
                               FP32 one, so steady flows do not accumulate a bias */
};

/* Shan-Chen variants (ExternalForce.py:55-145) */
enum { LBM_SC_NONE = 0, LBM_SC_C1 = 1, LBM_SC_C2 = 2 };

/* Pseudopotentials, EOS.py: SC_psi (:40-47) psi = rho0 (1 - exp(-rho/rho0));
 * the cubic equations of state (:5-95) psi = sqrt(|6/G (p(rho, T) - rho/3)|)
 * with T the node's temperature (lb_field.T) and eos = (a, b, R, G, omega,
 * T_cr):
 *   vdW  p = rho R T/(1 - b rho) - a rho^2                             (:50-58)
 *   RK   p = rho R T/(1 - b rho) - a rho^2 / (sqrt(T) (1 + b rho))      (:61-69)
 *   PR   p = rho R T/(1 - b rho) - a alpha(T) rho^2/(1 + 2 b rho - (b rho)^2),
 *        alpha = (1 + (0.37464 + 1.54226 omega - 0.269922 omega^2)
 *                     (1 - sqrt(T/T_cr)))^2                            (:72-83)
 *   CS   p = rho R T (1 + x + x^2 - x^3)/(1 - x)^3 - a rho^2, x = b rho/4 (:86-95) */
enum { LBM_PSI_SC = 0, LBM_PSI_VDW = 1, LBM_PSI_RK = 2, LBM_PSI_PR = 3, LBM_PSI_CS = 4 };

/* node code: one uint16 per node, compiled on the host from the ordered
 * boundary-condition list (BoundaryCondition.py:459-483).               */
#define LBM_CODE_BOUNCE_MASK 0x00FFu /* bit k-1 set: population k takes f2(x, opp k)   */
#define LBM_CODE_ACTIVE 0x0100u      /* node belongs to the fluid (collision) group     */
#define LBM_CODE_KEEP_MACRO 0x0200u  /* rho/u of this node are read by a node rule      */
#define LBM_CODE_NEEM_SHIFT 10       /* bits 10..13: node rule index + 1 (0 = none)     */
#define LBM_CODE_NEEM_MASK 0x3C00u
#define LBM_CODE_SC_SOLID 0x8000u    /* fluid_strategy is false here: a Shan-Chen      *
                                      * neighbour contributing adhesion, not cohesion  *
                                      * (ExternalForce.py:105-116)                      */
#define LBM_CODE_PULLED 0x4000u      /* node outside the fluid group whose populations  *
                                      * an active node pulls (a plain, not bounced,    *
                                      * population reads its unchanged f2)             */

/* Node-rule kinds: the boundary conditions that act on a node after its
 * populations were streamed or bounced.                                  */
enum {
  LBM_NEEM_VELOCITY = 1,     /* VelocityNEEM.apply, BoundaryCondition.py:149-166 */
  LBM_NEEM_PRESSURE = 2,     /* PressureNEEM.apply, :209-226                     */
  LBM_RULE_VELOCITY_BB = 3,  /* VelocityBB.apply, :85-102 (adds 6 w_k rho c_k.u) */
  LBM_RULE_OPEN_NEUMANN = 4, /* OpenNeumann.apply, :321-334 (unknowns := f(x_nb)) */
  LBM_RULE_PRESSURE_ABC = 5, /* PressureABC.apply, :190-205 (unknowns := -f + 2 w rho_value (...), *
                              * u_w = 1.5 u(x) - 0.5 u(x_nb))                                      */
  LBM_RULE_VELOCITY_EQ = 6,  /* VelocityEquilibrium.apply, :135-145 (all 9 := feq(u_w, rho_w))     */
  LBM_RULE_OPEN_CONVECTIVE1 = 7 /* OpenConvective1order.apply, :343-359: unknowns := (f + l f(x_nb)) / *
                                 * (1 + l), l = |u(x_nb)|, then density and velocity                  */
};

/* Where the relaxation time comes from.  The per-node state (lbm_fields.tau,
 * lbm_gfields.tau) holds BGK tau, TRT tau_sym or MRT diag[7] (= 1/tau); TRT
 * tau_antisym is always magic/(tau_sym - 0.5) + 0.5 (CollisionEngine.py:155,
 * 164) and the other MRT rates are diag = (0, s_bulk, 1, 0, 1, 0, 1, s, s)
 * (:301-318), so neither needs storage.                                    */
enum {
  LBM_TAU_CONST = 0,     /* uniform value from initTau (lbm_lattice.tau / s_shear) */
  LBM_TAU_FIELD = 1,     /* per-node state, read only                              */
  LBM_TAU_POWER_LAW = 2  /* shear_rate + relaxation_time each step, state updated  */
};
#define LBM_MAX_NEEM 15
#define LBM_MAX_NC 2

/* One node rule: VelocityNEEM (BoundaryCondition.py:149-166), PressureNEEM
 * (:209-226), VelocityBB (:85-102), OpenNeumann (:321-334), PressureABC
 * (:177-205) or VelocityEquilibrium (:132-145).                           */
typedef struct lbm_neem_rule {
  int32_t kind;      /* LBM_NEEM_VELOCITY | LBM_NEEM_PRESSURE | LBM_RULE_*   */
  int32_t direction; /* outward population index d in 1..4; x_nb = x - c_d */
  float ux, uy;      /* velocity_value (velocity NEEM, VelocityBB)          */
  float rho;         /* rho_value (pressure rules; PressureNEEM writes it to *
                      * total_rho only, PressureABC also imposes it)        */
  int32_t pad;
} lbm_neem_rule;

/* Lattice description and per-run physical parameters.  Parameters are read
 * at every call (the reference freezes Python scalars at first compile).   */
typedef struct lbm_lattice {
  int32_t nx, ny, nc;
  int32_t collision;  /* LBM_BGK | LBM_TRT | LBM_MRT                          */
  int32_t storage;    /* LBM_STORE_*                                          */
  int32_t tau_mode;   /* LBM_TAU_CONST | LBM_TAU_FIELD | LBM_TAU_POWER_LAW     */
  int32_t shan_chen;  /* LBM_SC_*                                             */
  int32_t n_neem;     /* number of rules in the lbm_neem_rule table           */
  int64_t plane_stride;
  /* relaxation (fp32 values exactly as the reference's initTau computes them)*/
  float tau[LBM_MAX_NC];       /* BGK tau / TRT tau_sym (constant-tau runs)   */
  float tau_anti[LBM_MAX_NC];  /* TRT tau_antisym (constant-tau runs)         */
  float magic[LBM_MAX_NC];     /* TRT magic parameter                         */
  float s_shear[LBM_MAX_NC];   /* MRT diag[7] = diag[8] = 1/tau               */
  float s_bulk[LBM_MAX_NC];    /* MRT diag[1] = 1/tau_bulk                    */
  float visc_m, visc_nm1;      /* nu_app = m * shear_rate^(n-1): m and (n - 1)
                                  computed in double, then rounded; Newtonian: (nu, 0) */
  float tau_min, tau_max;      /* Collision.tau_max_min bounds                */
  float gx, gy;                /* gravity_force (force density, lbm.py:220)   */
  float g_coh[LBM_MAX_NC * LBM_MAX_NC]; /* C1: g_coh[0]; C2: matrix row-major */
  float psi_rho0[LBM_MAX_NC];  /* SC_psi rho0 per component (EOS.py:40-47)    */
  int32_t psi_kind[LBM_MAX_NC]; /* LBM_PSI_*: the psi object of each component */
  float eos[LBM_MAX_NC][6];    /* cubic EOS (EOS.py:5-95): a, b, R, G, omega, T_cr */
  float g_adh[LBM_MAX_NC];     /* ShanChenForceC2 gadh (ExternalForce.py:105-116) */
  int32_t sc_adhesion;         /* 1: some node is not fluid for the Shan-Chen   *
                                * force (LBM_CODE_SC_SOLID / lbm_gfields.sc_fluid) */
  int32_t huang;               /* 1: HuangMRTCollision (CollisionEngine.py:453-515): *
                                * MRT with the source S Q, Q from the component's   *
                                * force and psi; needs LBM_MRT and Shan-Chen         */
  float huang_k1[LBM_MAX_NC];  /* k1                                               */
  float huang_k2[LBM_MAX_NC];  /* k2 = epsilon / (-8) - k1 (:462)                  */
  float huang_g[LBM_MAX_NC];   /* G of the component with itself (:463)            */
  uint32_t round_step;         /* steps the lattice has taken since its initial state *
                                * (mod 2^32): with FP16 / BF16 storage the stores of *
                                * step t are rounded stochastically, seeded by t and *
                                * the node, so a lattice rounds the same in split     *
                                * runs and in a batch; no counterpart in the          *
                                * reference (its fields are FP32)                     */
  int32_t rounding;            /* LBM_ROUND_*: FP16 / BF16 stores (ignored otherwise) */
  int32_t pad1;
} lbm_lattice;

/* Device buffers of one lattice (see the plane layout above). */
typedef struct lbm_fields {
  void* f[2];           /* population ping-pong, nc*9 planes each           */
  void* macro[2];       /* macro ping-pong, nc+3 planes each: double with  */
                        /* LBM_STORE_F64 storage, float otherwise           */
  void* tau;            /* nc planes, per-node relaxation state (double    */
                        /* with LBM_STORE_F64 storage, float otherwise)     */
  float* psi[2];        /* nc planes each, Shan-Chen psi ping-pong          */
  float* body_force;    /* nc*2 planes                                      */
  float* sc_force;      /* nc*2 planes                                      */
  float* shear_rate;    /* nc planes                                        */
  const uint16_t* code; /* one node code per node                           */
  void* tau2;           /* batches: a second relaxation-state buffer (as    */
                        /* tau) that lets power-law lattices take two steps */
                        /* per launch; NULL: one step per launch            */
  const float* temperature; /* 1 plane (lb_field.T): cubic-EOS psi only     */
} lbm_fields;

/* ---- version / errors ---------------------------------------------------*/
int lbm_abi_version(void);
const char* lbm_last_error(void);

/* ---- fused hot path -----------------------------------------------------*/
/* Runs `nsteps` calls of the cases' lbm_solve() (cases/Circular cylinder.py:
 * 182-193, cases/Power Law Fluid.py:160-173, cases/Bubble or two-phase
 * separation.py:197-212): density -> [update_psi -> psi halo -> cau_SCforce]
 * -> force_density -> velocity -> [shear_rate -> relaxation_time] -> collide,
 * with the boundary list of step n-1 (streaming, bounce-back, NEEM) fused in
 * front of step n as a pull.  Launch t reads f[p], macro[p] (p = (parity +
 * t) % 2) and writes f[1-p], macro[1-p]; the newest post-collision
 * populations end in f[(parity + nsteps) % 2].
 *   first_local = 1: f[parity] holds post-stream populations f (the state
 *   after init_LBM or after lbm_finish) and the first launch reads them in
 *   place; 0: f[parity] holds post-collision populations f2 and the first
 *   launch streams them.
 * The last launch also writes every node's rho/total_rho/u (and body force,
 * SC force, shear rate where enabled), as the reference's fields hold them
 * before the boundary pass.
 * Shan-Chen: launch t computes psi of its step (FP32 storage: on chip, per
 * tile) and leaves it for the active nodes in psi[1-p] only if it is one of
 * the last two launches; nodes outside the fluid group keep the psi the
 * planes hold (both planes should hold the same values there).            */
int lbm_run(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem,
            int32_t nsteps, int32_t parity, int32_t first_local, void* stream);

/* Same steps as lbm_run, two per launch where the lattice allows it
 * (temporal blocking: step n is kept in LDS and step n+1 computed from it,
 * halving the HBM traffic per step).  Lattices with node rules (n_neem > 0)
 * use a two-node ring around each tile; Shan-Chen and FP64 storage run one
 * step per launch.
 * A two-step launch flips the population / macro buffers once, so the index
 * of the buffer holding the newest f2 -- what lbm_finish takes -- is returned
 * in *out_parity.
 * tile_cols: 16, 32 or 64 tile columns (0: the default; lattices beyond
 * 32-bit buffer offsets always use 64); the block has one thread per node of
 * the ring-extended tile (256 to 1024).  With LBM_TAU_POWER_LAW the relaxation state is
 * double-buffered: tau_next is a second state buffer of the same size as
 * fields->tau (required, else one step per launch); the final state is left
 * in fields->tau.  Bit-identical to lbm_run.  No counterpart in the
 * reference: its steps are the same lbm_solve() sequence. */
int lbm_run_blocked(const lbm_lattice* lattice, const lbm_fields* fields, const lbm_neem_rule* neem,
                    int32_t nsteps, int32_t parity, int32_t first_local, int32_t tile_cols, void* tau_next,
                    int32_t* out_parity, void* stream);
/* lbm_run_blocked without the tau_next argument, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The second relaxation-state buffer of two-step
 * power-law runs is a fixed workspace of 2^22 floats (16 MiB) in device
 * memory that every GPU the library runs on holds until the process ends,
 * so nc * plane_stride must not exceed 4194304 (a 2048 x 2048 lattice with
 * one component, 2048 x 1024 with two); larger power-law lattices are
 * refused with LBM_E_INVALID.  The call makes no hipMalloc, hipFree or other
 * allocating call.  All calls on one GPU share its workspace and nothing in
 * it needs to start at zero: calls that overlap on one GPU (different
 * streams or host threads) race and must be serialised by the caller.
 * Results are those of lbm_run_blocked bit for bit.                       */
int lbm_run_blocked_noalloc(const lbm_lattice* lattice, const lbm_fields* fields, const lbm_neem_rule* neem,
                            int32_t nsteps, int32_t parity, int32_t first_local, int32_t tile_cols,
                            int32_t* out_parity, void* stream);
/* Rows of a two-step tile with `ring` = 1 or 2 and `tile_cols` = 16, 32 or
 * 64 columns (0: the library's default width).                            */
int lbm_tile_rows(int32_t ring, int32_t tile_cols);

/* The reference's boundary pass alone (BoundaryEngine.apply_boundary_conditions,
 * BoundaryCondition.py:481-483) applied to f2 = f[parity]: writes the
 * post-stream f into f[1-parity] and applies the NEEM overrides of vel / rho /
 * total_rho (BoundaryCondition.py:161-163, 222-223) to macro[parity].        */
int lbm_finish(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem,
               int32_t parity, void* stream);

/* ---- batches of independent lattices --------------------------------------
 * Many lattices (each with its own lbm_lattice, lbm_fields and NEEM rules;
 * same nc and storage, no Shan-Chen) advance together in one persistent
 * launch per step over their 256-node tiles.  No counterpart in the
 * reference, which steps one lattice per kernel sequence; each lattice still
 * follows lbm_run's semantics exactly.  The batch object owns a device copy
 * of the descriptors, the tile table and residual accumulators.          */
typedef struct lbm_batch lbm_batch;
/* neem: count x LBM_MAX_NEEM rules (lattice l's rules start at l*LBM_MAX_NEEM) */
int lbm_batch_create(const lbm_lattice* lats, const lbm_fields* flds, const lbm_neem_rule* neem, int32_t count,
                     lbm_batch** out);
/* Replace lattice `index` (its buffers / parameters); takes effect at commit. */
int lbm_batch_update(lbm_batch* b, int32_t index, const lbm_lattice* lat, const lbm_fields* fld,
                     const lbm_neem_rule* neem);
/* active: include in the tile list; start_step: launch index of a run at which
 * this lattice's populations are post-stream f read in place (-1: never);
 * vel_old: 2 planes for lbm_batch_residual (may be NULL).                  */
int lbm_batch_set(lbm_batch* b, int32_t index, int32_t active, int32_t start_step, float* vel_old);
/* The step count of lattice `index` (lbm_lattice.round_step of a batch: kept
 * in device memory, advanced by every run and graph replay of the batch, 0
 * at lbm_batch_create); set it when a slot takes a new lattice.  Ordered on
 * `stream`.  No counterpart in the reference (16-bit storage).             */
int lbm_batch_set_round_step(lbm_batch* b, int32_t index, uint32_t step, void* stream);
/* Rebuild the tile list and upload descriptors (synchronous on `stream`).  A
 * graph captured earlier stays valid while the tile count fits the table
 * capacity of its capture (lbm_batch_capacity).                           */
int lbm_batch_commit(lbm_batch* b, void* stream);
int lbm_batch_tiles(const lbm_batch* b);
int lbm_batch_capacity(const lbm_batch* b);
/* Reserve room for `tiles` tiles in the one-step tile table, or `tiles` in
 * the two-step table (reserve2), before capturing a graph: after a capture a
 * commit that needs more room fails instead of moving the tables.  reserve2
 * also shrinks the two-step table (to no fewer entries than the current
 * tile count) while no graph holds it: a captured graph launches one block
 * per entry of the capacity.                                               */
int lbm_batch_reserve(lbm_batch* b, int32_t tiles);
int lbm_batch_reserve2(lbm_batch* b, int32_t tiles);
/* nsteps launches (launch t reads parity (parity + t) % 2; the last writes
 * every node's macro fields).  grid_blocks: persistent grid, multiple of 8
 * (0 = 2048).                                                              */
int lbm_batch_run(lbm_batch* b, int32_t nsteps, int32_t parity, int32_t grid_blocks, void* stream);
/* lbm_finish for every active lattice of the batch. */
int lbm_batch_finish(lbm_batch* b, int32_t parity, int32_t grid_blocks, void* stream);
/* Capture lbm_batch_run(b, nsteps, parity, grid_blocks) into an executable
 * graph (stream must not be NULL); replay with lbm_graph_launch.          */
int lbm_batch_graph(lbm_batch* b, int32_t nsteps, int32_t parity, int32_t grid_blocks, void* stream,
                    void** graph_exec);
/* Two steps per launch for the whole batch where it allows it (as
 * lbm_run_blocked: every active lattice without power law, storage other
 * than FP64).  lbm_batch_mode reports what the last commit chose: 0 = one
 * step per launch (also when a power-law lattice has no fields.tau2);
 * 1 = two-step launches with a one-node ring (no node
 * rules); 2 = a two-node ring (node rules); 3 = node rules with a one-node
 * ring: each lattice's tile grid is shifted (found at commit by scanning the
 * rule nodes) so that every NEEM / OpenNeumann node shares its tile with the
 * interior neighbour it reads; plus 16 when no active lattice has a body
 * force (one-component batches in 32- or 64-column tiles then run kernels
 * built without the force terms).  lbm_batch_ring(b, 2) forces mode 2 instead of
 * 3 (0 = automatic); lbm_batch_tiles2 is the two-step tile count.  A two-step launch flips the
 * buffers once; the parity of the buffer with the newest f2 is returned in
 * *out_parity.  grid_blocks applies to the one-step launches only; a
 * two-step launch has one block per tile (per entry of the table's capacity
 * in a graph).  The graph variant captures the same launch sequence; it
 * stays valid while the mode does not change.  No counterpart in the
 * reference. */
int lbm_batch_mode(const lbm_batch* b);
int lbm_batch_ring(lbm_batch* b, int32_t ring);
int lbm_batch_tiles2(const lbm_batch* b);
/* The caller has destroyed every graph captured from this batch: its tables
 * and tile shape may change again (a capture pins them).                  */
int lbm_batch_release(lbm_batch* b);
/* Tile columns (0: default, 16, 32 or 64) of the batch's two-step launches from
 * the next commit on; refused after a graph capture.                      */
int lbm_batch_tile_cols(lbm_batch* b, int32_t tile_cols);
int lbm_batch_run_blocked(lbm_batch* b, int32_t nsteps, int32_t parity, int32_t grid_blocks, int32_t* out_parity,
                          void* stream);
int lbm_batch_graph_blocked(lbm_batch* b, int32_t nsteps, int32_t parity, int32_t grid_blocks, void* stream,
                            void** graph_exec, int32_t* out_parity);
int lbm_graph_launch(void* graph_exec, void* stream);
int lbm_graph_destroy(void* graph_exec);
/* out[2l] = sum |u - u_old|^2, out[2l+1] = sum |u|^2 over lattice l's active
 * nodes (u from macro[parity]); then u_old = u.  Synchronous.              */
int lbm_batch_residual(lbm_batch* b, int32_t parity, double* out, void* stream);
int lbm_batch_destroy(lbm_batch* b);

/* ---- reductions (FP64) --------------------------------------------------*/
/* out[c] = sum over active nodes of rho_c; out[nc], out[nc+1] = sum of
 * total_rho*u (momentum); out[nc+2] = number of non-finite values among
 * rho/u.  Reads macro planes; synchronous on `stream`.                     */
int lbm_reduce_macro(const lbm_lattice* lat, const float* macro, const uint16_t* code,
                     double* out, void* stream);
/* Residual of cases/Power Law Fluid.py:183-192: returns sum |u - u_old|^2
 * over all nodes and then copies u into u_old (2 planes each).            */
int lbm_residual(int64_t n_nodes, int64_t plane_stride, const float* vel, float* vel_old,
                 double* out, void* stream);

/* ---- granular kernels: one per reference @ti.kernel ----------------------
 * Fields use the plane layout; `group` is an int32 (count, 2) list of (i, j)
 * in device memory.  These reproduce the unfused reference sequence.       */
typedef struct lbm_gfields {
  float* f;            /* post-stream populations (lb_field.f)   */
  float* f2;           /* post-collision populations (lb_field.f2) */
  float* rho;          /* nc planes */
  float* total_rho;    /* 1 plane   */
  float* vel;          /* 2 planes  */
  float* body_force;   /* nc*2 planes */
  float* sc_force;     /* nc*2 planes */
  float* shear_rate;   /* nc planes */
  float* pressure;     /* nc planes */
  float* total_pressure; /* 1 plane */
  float* temperature;  /* 1 plane (lb_field.T) */
  float* psi;          /* nc planes of (NX+2) x (NY+2), psi(-1,-1) at offset 0 */
  const int32_t* neighbor; /* (NX, NY, 9, 2) int32, lbm.py:50 layout */
  float* tau;          /* nc planes: BGK tau / TRT tau_sym / MRT diag[7] */
  const uint8_t* sc_fluid; /* (NX+2) x (NY+2) halo grid, offset -1: 1 where the *
                            * ShanChenForceC2 fluid_strategy is true (read when *
                            * lbm_lattice.sc_adhesion is set)                   */
} lbm_gfields;

/* LBField.init_LBM, lbm.py:119-125 */
int lbm_g_init_lbm(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count, void* stream);
/* MacroscopicEngine.density, lbm.py:153-164 */
int lbm_g_density(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count, void* stream);
/* MacroscopicEngine.pressure0/1/2, lbm.py:166-212 (variant from lat->shan_chen and nc) */
int lbm_g_pressure(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count, void* stream);
/* MacroscopicEngine.force_density, lbm.py:215-220 */
int lbm_g_force_density(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count, void* stream);
/* MacroscopicEngine.velocity, lbm.py:223-237 */
int lbm_g_velocity(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count, void* stream);
/* {BGK,TRT,MRT}Collision.shear_rate, CollisionEngine.py:80-92, 166-178, 332-377 */
int lbm_g_shear_rate(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count, void* stream);
/* {BGK,TRT,MRT}Collision.relaxation_time, CollisionEngine.py:72-78, 157-164, 322-329 */
int lbm_g_relaxation_time(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count, void* stream);
/* {BGK,TRT,MRT}Collision.apply, CollisionEngine.py:96-110, 223-245, 424-451 */
int lbm_g_collide(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count, void* stream);

/* boundary kinds for lbm_g_boundary */
enum {
  LBM_BC_PERIODIC = 1,       /* PeriodicAllBoundary.apply, BoundaryCondition.py:424-432 */
  LBM_BC_INSIDE = 2,         /* InsideBoundary.apply, :440-449 (-1 entries are skipped) */
  LBM_BC_BOUNCEBACK = 3,     /* BounceBackWall.apply, :244-259 */
  LBM_BC_VELOCITY_NEEM = 4,  /* VelocityNEEM.apply, :154-166 */
  LBM_BC_PRESSURE_NEEM = 5,  /* PressureNEEM.apply, :215-226 */
  LBM_BC_VELOCITY_BB = 6,    /* VelocityBB.apply, :85-102 (additive, rule.ux/uy, direction) */
  LBM_BC_OPEN_NEUMANN = 7,   /* OpenNeumann.apply, :321-334 (copy unknowns, then density and velocity) */
  LBM_BC_PRESSURE_ABC = 8,   /* PressureABC.apply, :190-205 (rule.rho, direction)               */
  LBM_BC_VELOCITY_EQ = 9,    /* VelocityEquilibrium.apply, :135-145 (rule.ux/uy, direction)     */
  LBM_BC_OPEN_CONVECTIVE1 = 10, /* OpenConvective1order.apply, :343-359 (direction)                */
  LBM_BC_OPEN_EXTRAPOLATION = 11 /* OpenExtrapolation.apply, :399-413: unknowns := 2 f(x - c_d)   *
                                  * - f(x - 2 c_d) (direction)                                    */
};
int lbm_g_boundary(const lbm_lattice* lat, const lbm_gfields* g, int32_t kind, const lbm_neem_rule* rule,
                   const int32_t* group, int32_t count, void* stream);

/* ShanChenForceC1/C2.update_psi, ExternalForce.py:70-75, 114-120 */
int lbm_g_sc_update_psi(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count, void* stream);
/* PeriodicPsi_LR (which & 1) then PeriodicPsi_BT (which & 2), SCBoundaryCondition.py:126-144 */
int lbm_g_sc_periodic_psi(const lbm_lattice* lat, const lbm_gfields* g, int32_t which, void* stream);
/* ShanChenForceC1/C2.cau_SCforce, ExternalForce.py:78-91, 124-145 */
int lbm_g_sc_force(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count, void* stream);

/* psi boundaries other than the periodic pair, SCBoundaryCondition.py:9-123.
 * The group lists halo-grid nodes (i, j) in [-1, NX] x [-1, NY]; the
 * interior neighbour is x - c_direction.                                   */
enum {
  LBM_PSI_BC_NEUMANN = 1,       /* OpenNeumannPsi :19-33: psi_c(x) := psi_c(x - c_d)          */
  LBM_PSI_BC_CONVECTIVE1 = 2,   /* OpenConvective1orderPsi :36-58                              */
  LBM_PSI_BC_CONVECTIVE2 = 3,   /* OpenConvective2orderPsi :61-89                              */
  LBM_PSI_BC_EXTRAPOLATION = 4, /* OpenExtrapolationPsi :92-107: 2 psi(x - c_d) - psi(x - 2c_d) */
  LBM_PSI_BC_FIXED = 5          /* FixedPsi :110-123: psi_0(x) := value (component 0 only)     */
};
/* The convective kinds average the velocity of the group's interior
 * neighbours (one block reduces the group, so groups are boundary lines). */
int lbm_g_sc_psi_boundary(const lbm_lattice* lat, const lbm_gfields* g, int32_t kind, int32_t direction,
                          float value, const int32_t* group, int32_t count, void* stream);

/* ---- batched Shan-Chen lattices -----------------------------------------
 * Many independent Shan-Chen lattices (ShanChenForceC1/C2, ExternalForce.py:
 * 55-145, each with its own lbm_solve() of cases/Bubble or two-phase
 * separation.py:197-212) advanced together: one launch per pair of steps
 * and collision model runs the psi-in-LDS two-step tile of every lattice
 * (one launch per step where a lattice has node rules or a power law, and
 * for a run's odd last step).  Each lattice keeps its own
 * buffers; lattice l's rules are neem[l * LBM_MAX_NEEM ...].  FP32 storage,
 * the same num_components for all.  parity_offset[l] (0 / 1, NULL: all 0)
 * is lattice l's ping-pong phase against the run's `parity`: a launch reads
 * pop[launch parity ^ parity_offset[l]].  Every lattice's fields equal its
 * lbm_run_blocked alone bit for bit.                                           */
typedef struct lbm_sc_batch lbm_sc_batch;
int lbm_sc_batch_create(const lbm_lattice* lats, const lbm_fields* flds, const lbm_neem_rule* neem,
                        const int32_t* parity_offset, int32_t count, lbm_sc_batch** out);
/* nsteps steps of every lattice (first_local: the first step reads each
 * lattice's post-stream f in place, as lbm_run's); *out_parity: the run's
 * parity after the last launch (a launch of two steps flips it once), so
 * lattice l's newest f2 is pop[*out_parity ^ parity_offset[l]]            */
int lbm_sc_batch_run(lbm_sc_batch* b, int32_t nsteps, int32_t parity, int32_t first_local, int32_t* out_parity,
                     void* stream);
int lbm_sc_batch_tiles(const lbm_sc_batch* b);
int lbm_sc_batch_destroy(lbm_sc_batch* b);

/* ---- immersed-boundary particles (ibm.py) --------------------------------
 * Circular particles (Sphere, ibm.py:17-26) carrying a ring of Lagrangian
 * markers (IBMNodes, ibm.py:8-14), coupled to the lattice by multi-direct
 * forcing (IBEngine, ibm.py:166-253).  Sphere fields are arrays over the
 * max_number capacity; marker fields are arrays over max_number *
 * max_num_ib with marker (m, n) at m * max_num_ib + n; vectors are two
 * consecutive planes (x, then y) of such an array.                         */
typedef struct lbm_ib_particles {
  int32_t max_number;      /* SphereIB Max_Number                              */
  int32_t max_num_ib;      /* SphereIB Max_NumIB                               */
  int32_t num;             /* SphereIB.num[None]: spheres in use, <= max_number */
  float numib_cof;         /* SphereIB.NUMIB_cof (marker spacing)              */
  float* pos;              /* Sphere.pos, 2 planes of max_number               */
  float* vel;              /* Sphere.vel                                       */
  float* force;            /* Sphere.force                                     */
  float* torque;           /* Sphere.torque, 1 plane                           */
  float* radius;           /* Sphere.radius                                    */
  float* angle;            /* Sphere.angle                                     */
  int32_t* num_ib;         /* Sphere.NumIB                                     */
  float* node_pos;         /* IBMNodes.pos, 2 planes of max_number*max_num_ib  */
  float* node_vel;         /* IBMNodes.vel                                     */
  float* node_vel_desired; /* IBMNodes.vel_desired                             */
  float* node_force;       /* IBMNodes.force                                   */
  float* node_force_temp;  /* IBMNodes.force_temp                              */
  int32_t* group;          /* IBNodeInsiderGroup: (max_number*max_num_ib, 2)   *
                            * int32 (m, n), the first *count entries in use     */
  int32_t* count;          /* IBInsiderCount[None], one int32                  */
  int32_t* work;           /* lbm_ib_work_ints(max_number, max_num_ib) int32 of *
                            * scratch, zero-filled before first use: marker      *
                            * ranks, per-sphere counts and offsets of the last   *
                            * lbm_ib_update (read by lbm_ib_particle_force)      */
} lbm_ib_particles;

/* Size of lbm_ib_particles.work in int32 (0 for a bad capacity). */
int64_t lbm_ib_work_ints(int32_t max_number, int32_t max_num_ib);
/* SphereIB.init_IB_nodes, ibm.py:57-67: NumIB = round(2 pi r / numib_cof)
 * markers on each sphere's circle (markers beyond max_num_ib are not set). */
int lbm_ib_init_nodes(const lbm_ib_particles* p, void* stream);
/* IBDEMCouplerEngine.SphereIBUpdate, ibm.py:120-141 (reads lat->nx, ny):
 * moves the markers of spheres inside or crossing the lattice
 * (computer_IB_nodes_pos, :68-77), sets their desired velocity, and lists
 * the markers in use in (m, n) order.                                      */
int lbm_ib_update(const lbm_lattice* lat, const lbm_ib_particles* p, void* stream);
/* SphereIB.reset_force, ibm.py:80-83 */
int lbm_ib_reset_force(const lbm_ib_particles* p, void* stream);
/* IBEngine.InterpolateBodyVelocities, ibm.py:201-223 (reads g->vel) */
int lbm_ib_interpolate(const lbm_lattice* lat, const lbm_gfields* g, const lbm_ib_particles* p, void* stream);
/* IBEngine.SpreadBodyForces, ibm.py:225-244 (reads g->rho, g->total_rho;
 * adds to g->vel and g->body_force with float atomics)                    */
int lbm_ib_spread(const lbm_lattice* lat, const lbm_gfields* g, const lbm_ib_particles* p, void* stream);
/* IBEngine.MultiDirectForcingLoop, ibm.py:248-253: reset_force, then
 * `iterations` x (interpolate, spread).                                    */
int lbm_ib_mdf(const lbm_lattice* lat, const lbm_gfields* g, const lbm_ib_particles* p, int32_t iterations,
               void* stream);
/* nsteps lbm_solve() steps of a lattice with immersed-boundary particles,
 * IBEngine.MultiDirectForcingLoop (`iterations`) between the velocity and
 * the collision of every step (ibm.py:248-253), fused: per step the
 * stencil nodes of the marker list get the step's density and velocity in
 * `scratch` (vel 2, rho nc, total_rho 1, body_force 2*nc planes of
 * lat->plane_stride), the forcing loop runs on them, one fused step runs
 * every node, and the stencil nodes are stepped again with the spread force.
 * The marker list is the caller's last lbm_ib_update (particles do not move
 * within a call).  FP32 storage, no node rules, Shan-Chen or power law.
 * parity / first_local / *out_parity as lbm_run_blocked.                  */
int lbm_ib_run(const lbm_lattice* lat, const lbm_fields* fld, const lbm_ib_particles* p, const lbm_gfields* scratch,
               int32_t iterations, int32_t nsteps, int32_t parity, int32_t first_local, int32_t* out_parity,
               void* stream);
/* IBEngine.ComputeParticleForce, ibm.py:171-185: each sphere with markers
 * in the list of the last lbm_ib_update gets force = sum of their forces and
 * torque = sum (F_x r_y - F_y r_x), r = marker - centre; other spheres keep
 * theirs.  (The reference resets the sums inside the marker loop, a race.) */
int lbm_ib_particle_force(const lbm_ib_particles* p, void* stream);

/* ---- memory-bandwidth probe -----------------------------------------------*/
/* Plain fp32 copy of n elements (the achievable-bandwidth reference used to
 * scale HBM counters and to quote a measured copy ceiling).               */
int lbm_copy_f32(const float* src, float* dst, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TAICHI_LBM_AMD_LBM_ABI_H */
