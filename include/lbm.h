/*
 * lbm.h -- C ABI of the MI355X lattice-Boltzmann engine (liblbm.so).
 *
 * The reference drives its solver through WebGPU from JavaScript.  Each entry
 * point below replaces one group of those WebGPU calls; the reference call
 * site it stands in for is cited beside it.  Plain C types only: opaque
 * handles, POD descriptors, caller-owned host arrays copied in and out.
 * Every function returns 0 on success or a negative LBM_E_* code, and
 * lbm_last_error() then returns a message for the calling thread.
 *
 * Threading: one host thread drives a given solve handle at a time (the
 * reference has one WebGPU queue per page).  Different handles may be driven
 * from different threads.
 */
#ifndef LBM_H
#define LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 7

enum {
  LBM_OK = 0,
  LBM_E_INVALID = -1,   /* bad argument or descriptor */
  LBM_E_DEVICE = -2,    /* HIP runtime error */
  LBM_E_NOMEM = -3,     /* device allocation failed / budget exceeded */
  LBM_E_REFUSED = -4,   /* configuration the reference refuses */
  LBM_E_STATE = -5      /* call out of order (e.g. step before init) */
};

enum { LBM_BGK = 0, LBM_TRT = 1 };

/* Population storage in HBM.  Collision is always computed in FP32.
 * F16/BF16 store the shifted value f_i - w_i (the reference's f-pack mode 2,
 * common_fpack.wgsl:91-121, extended from D2Q9 to D3Q19/D3Q27). */
enum { LBM_STORE_F32 = 0, LBM_STORE_F16 = 1, LBM_STORE_BF16 = 2 };

enum { LBM_BODY_NONE = 0, LBM_BODY_SPHERE = 1, LBM_BODY_CIRCLE = 2 };

/* Pinned rigid body (d3-body.mjs BODY_FIELDS:318 / CardState
 * common_geometry.wgsl:8-35, restricted to what a pinned body reads). */
typedef struct lbm_body {
  int32_t shape;        /* LBM_BODY_* ; SPHERE for Q19/Q27, CIRCLE for Q9   */
  float radius;         /* a                                                  */
  float centre[3];      /* cx, cy, cz in lattice cells                        */
} lbm_body;

/* Solve descriptor: the reference's pipeline override constants
 * (common_d3_step.wgsl:94-179, lbm_step.wgsl:33-126) plus the grid. */
typedef struct lbm_desc {
  int32_t abi_version;  /* LBM_ABI_VERSION                                    */
  int32_t q;            /* 9 (D2Q9), 19 (D3Q19) or 27 (D3Q27)                  */
  int32_t nx, ny, nz;   /* nz must be 1 for Q = 9                             */
  int32_t collision;    /* LBM_BGK or LBM_TRT                                 */
  float omega;          /* 1/tau (even rate; OMEGA)                           */
  float omega_minus;    /* TRT odd rate (OMEGA_M, plans/TRT.md:352-367)       */
  float force[3];       /* uniform Guo body force (FORCE_X/Y/Z)               */
  int32_t walls[3];     /* halfway no-slip walls per axis (WALL_X/Y/Z)        */
  float wall_u[2];      /* x-velocity of the y=0 / y=ny-1 walls (WALL_U0/U1)  */
  float sponge_width;   /* SPONGE_W; <= 0 disables                            */
  int32_t sponge_axes[3]; /* SPONGE_AX/AY/AZ                                  */
  float sponge_u[3];    /* SPONGE_UX/UY/UZ                                    */
  lbm_body body;        /* pinned body with halfway bounce-back               */
  int32_t solid_eq;     /* SOLID_EQ: hold the body interior at feq(1, 0)      */
  int32_t storage;      /* LBM_STORE_*                                        */
  int32_t autotune;     /* 1: time candidate tile shapes on first step        */
  int32_t diffuse;      /* the body's coupling.  0: halfway bounce-back with
                         * momentum exchange (?bounceback, USE_BOUNCEBACK).
                         * 1: the diffuse penalty coupling, the reference's
                         * default (common_d3_step.wgsl:288-300): no bounce-
                         * back, every cell takes the Guo force rho chi
                         * (0 - u*), u* = m / rho, chi = 1/2 (1 - tanh(phi /
                         * chi_eps)), and the body feels rho chi u* summed
                         * over the cells with chi >= 1e-6
                         * (common_d3_force.wgsl:242-268).  SOLID_EQ does not
                         * apply.  Standalone solves (not the serve loop).   */
  float chi_eps;        /* CHI_EPS (main-3d.js:1026); <= 0: 1.5               */
} lbm_desc;

typedef struct lbm_solve lbm_solve;

/* Library / device. */
const char* lbm_last_error(void);
int lbm_abi_version(void);
int lbm_device_info(char* name, size_t name_len, double* hbm_bytes, int32_t* cu_count);

/* Creates a dense solve: device buffers for two population lattices and the
 * macroscopic field, plus the solid classification of the body.
 * Replaces main-3d.js:752-807 (createBuffer) and :989-1048
 * (createComputePipelineAsync with stepConstants), main-tgv.js:218-235. */
int lbm_solve_create(const lbm_desc* desc, lbm_solve** out);
int lbm_solve_destroy(lbm_solve* s);

/* Uploads [rho, ux, uy, uz] per cell (float32, 4*ncells) and builds the
 * equilibrium populations: writeBuffer(mac) + initEq,
 * main-3d.js:2973-2994 / common_d3_step.wgsl:206-222. */
int lbm_upload_macro(lbm_solve* s, const float* mac4);

/* Populations in the f32 plane-major wire format f[i*ncells + cell]
 * (f-pack.mjs:6-14); packed to the storage format on the way in and
 * unpacked on the way out. Replaces writeF / snapshot readback. */
int lbm_upload_f(lbm_solve* s, const float* f);
int lbm_read_f(lbm_solve* s, float* f);

/* Exact checkpoint of a solve: this header, then the current lattice's
 * populations in the solve's own storage format (shifted f - w, f32 / f16 /
 * bf16 bits), plane-major.  The wire format above is an interchange format:
 * for fp32 storage its f = g + w rounds away low bits of g, so a run resumed
 * from it drifts from the uninterrupted one.  A run resumed from this state
 * with the same descriptor gives the fields and force records of the
 * uninterrupted run bit for bit (the 16-bit dither counter travels with it).
 * The reference's snapshots are the wire format (f-pack.mjs:6-14,
 * debugSnapshotSave main-amr.js:2017); this exactness has no counterpart
 * there.  All integers little-endian. */
typedef struct lbm_state_header {
  char magic[8];         /* "LBMSTATE"                                          */
  uint32_t version;      /* 1                                                   */
  int32_t q, nx, ny, nz; /* lattice and grid                                    */
  int32_t storage;       /* LBM_STORE_*                                         */
  uint32_t salt;         /* dither counter of the next step                     */
  int64_t step;          /* steps done                                          */
  uint64_t payload_bytes;/* q * ncells * element bytes that follow              */
  uint32_t body_bytes;   /* 0, or the size of the moving rigid body's record
                            (FP64 state and parameters) after the populations  */
  uint8_t reserved[4];
} lbm_state_header;

/* Bytes of this solve's state (header + populations [+ rigid body]).  A
 * state with a rigid body loads only into a solve with a body attached, whose
 * record it replaces; one without loads only into a solve without. */
int lbm_state_size(lbm_solve* s, size_t* bytes);
/* Writes the state after the work enqueued so far; buf holds >= state_size. */
int lbm_save_state(lbm_solve* s, void* buf, size_t bytes);
/* Replaces the populations, step count and dither counter by a state saved
 * from a solve of the same q, grid and storage (LBM_E_INVALID otherwise).
 * The force history restarts; the macroscopic field is written by the next
 * step. */
int lbm_load_state(lbm_solve* s, const void* buf, size_t bytes);

/* Advances n steps (force reduction then fused stream-collide per step) and
 * leaves the macroscopic field of the last step readable.  Replaces
 * encodeSteps / debugStepSync, main-3d.js:3122-3171 and :3506,
 * main-tgv.js:245-265.  Asynchronous with respect to the host. */
int lbm_step(lbm_solve* s, int32_t n);

/* lbm_step(n) bracketed by HIP events on the solve's own stream; returns the
 * device time of the n steps in milliseconds. */
int lbm_step_timed(lbm_solve* s, int32_t n, double* ms);

/* Blocks until all queued work of the solve is complete. */
int lbm_sync(lbm_solve* s);

/* [rho, ux, uy, uz] per cell, float32: readMacro, main-3d.js:3539. */
int lbm_read_macro(lbm_solve* s, float* mac4);

/* Momentum-exchange force and torque on the pinned body, FP64, one record
 * [fx, fy, fz, tx, ty, tz] per step since the last lbm_upload_* (oldest
 * first).  *count receives the number of records written (<= max_records).
 * Replaces the force atomics + drain, common_d3_force.wgsl:171-283 and
 * d3_physics.wgsl:56-75, read by readBody, main-3d.js:4721-4738. */
int lbm_read_forces(lbm_solve* s, double* out, int32_t max_records, int32_t* count);

/* FP64 totals over the current populations: [mass, px, py, pz]
 * (the extensive budget of readStats, main-3d.js:3610-3650). */
int lbm_read_totals(lbm_solve* s, double* out4);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the totals of lbm_read_totals for a lattice the caller
 * holds in device memory, written to the caller's device array out4
 * ([mass, px, py, pz], FP64) and enqueued on `stream` (a hipStream_t; null:
 * the null stream) without waiting.  The lattice is in the engine's device
 * layout: shifted populations g = f - w_i in the storage format (LBM_STORE_*),
 * cells in blocks of 64, each block holding its q populations as q runs of
 * 64 elements (population i of cell c at element ((c / 64) q + i) 64 +
 * c % 64).  For the lattice a solve holds, the four sums equal
 * lbm_read_totals' bit for bit.  Allocates nothing: the per-block partial
 * sums and the completion counter live in one fixed device array of the
 * library, LBM_NOALLOC_SCRATCH_BYTES on each GPU it runs on, held until the
 * process ends.  ncells from 1 to LBM_NOALLOC_MAX_CELLS (LBM_E_INVALID
 * beyond).  Every call on a GPU shares that array: calls that overlap on one
 * GPU (from two streams or two threads) race. */
#define LBM_NOALLOC_SCRATCH_BYTES 33024
#define LBM_NOALLOC_MAX_CELLS 2147483648LL
int lbm_totals_noalloc(int32_t q, int32_t storage, int64_t ncells, const void* lattice, double* out4,
                       void* stream);

/* Solid classification, one byte per cell: bit 0 = inside the body
 * (phi < 0), bit 1 = phi < 2 (a cell that can have a solid neighbour). */
int lbm_read_classes(lbm_solve* s, uint8_t* out);

/* Current step count, tile shape, tile flags (as lbm_set_tile takes them)
 * and device bytes held by the solve. */
int lbm_solve_info(lbm_solve* s, int64_t* step, int32_t* block_x, int32_t* block_y,
                   int32_t* flags, int64_t* device_bytes);

/* Overrides the autotuned tile shape: block_x * block_y threads (a multiple
 * of 64, at most 256); flags bit 0 maps consecutive tiles onto one XCD, bit 1
 * selects non-temporal loads and stores, bit 2 makes tiles linear (a block
 * steps block_x * block_y consecutive cells of the x-fastest index instead of
 * a block_x by block_y patch of one z plane), bits 4-7 cap residency at that
 * many waves per SIMD (0: no cap; 4 is available for every shape). */
int lbm_set_tile(lbm_solve* s, int32_t block_x, int32_t block_y, int32_t flags);

/* Measurement hook, no counterpart in the reference (its spike page times
 * passes with timestamp queries, main-3d-spike.js:229-286): launches the
 * fused stream-collide kernel n times from the current lattice into the other
 * one without swapping (the state does not advance) and returns the mean
 * device time of one launch in milliseconds, bracketed by HIP events on the
 * solve's stream. */
int lbm_time_step_kernel(lbm_solve* s, int32_t n, double* ms_per_launch);

/* Measurement hook, no counterpart in the reference: a 16-byte streaming
 * copy of `bytes` between two fresh device buffers, repeated `reps` times;
 * returns the achieved read+write bandwidth in GB/s (the practical HBM
 * ceiling the roofline fraction is read against) and the mean time of one
 * copy. */
int lbm_probe_copy_bandwidth(size_t bytes, int32_t reps, double* gbs, double* ms_per_copy);

/* ------------------------------------------------------------------------
 * Moving rigid body of a dense 3D solve (sharp bounce-back coupling).  The
 * body state is FP64 on the device and advances every step, in the
 * reference's order (main-3d.js:3109-3148): the fluid force and torque on
 * the pre-stream populations at the current pose (momentum exchange with
 * Ladd's moving-wall term at the local wall density, plus the swept-cell
 * term, common_d3_force.wgsl:171-222), the 6-DOF integrator
 * (d3_physics.wgsl:56-158, d3-body.mjs:234-297), the solid classification
 * at the new pose, then the fluid step at the new pose (moving-wall
 * bounce-back, the interior held at feq(1, u_body), the sponge in window
 * coordinates, common_d3_step.wgsl:224-400).  The force records go to the
 * solve's force history (lbm_read_forces).
 * ------------------------------------------------------------------------ */
enum { LBM_SHAPE_SPHERE = 0, LBM_SHAPE_SPHEROID = 1, LBM_SHAPE_ROUNDBOX = 2 };  /* d3-body.mjs:36 SHAPE */

/* Replaces packBodyState / makeBodyState (d3-body.mjs:299-354) plus the
 * physics pipeline's override constants (main-3d.js:1060-1070). */
typedef struct lbm_rigid_body {
  int32_t shape;          /* LBM_SHAPE_*                                          */
  int32_t pinned;         /* 1: force measured, pose held                         */
  double a, b, c, r;      /* sphere a; spheroid semi-axes (a, a, c); round box
                             half-extents (a, b, c) and corner radius r           */
  double density;         /* mass = density * volume                              */
  double x[3];            /* centre of mass, cells                                */
  double q[4];            /* orientation [w, x, y, z], body to world              */
  double v[3];            /* velocity, cells per step                             */
  double omega[3];        /* angular velocity (world), radians per step           */
  double gravity[3];      /* acceleration (GX/GY/GZ), buoyancy-corrected          */
  double v_max, o_max;    /* magnitude clamps on |v| and |omega|                  */
  int32_t planar;         /* PLANAR: no z translation, rotation about z only      */
  int32_t no_fluid_force; /* NO_FLUID_FORCE: force measured, never applied         */
  int32_t swept_force;    /* SWEPT_FORCE: the swept-cell term (1 in the reference)  */
  int32_t rho_w_local;    /* RHO_W_LOCAL: local wall density (1 in the reference) */
  int32_t window[3];      /* moving-window axes: the body wraps, the sponge follows */
  int32_t reserved;
  double window_anchor[3]; /* WIN_AX/AY/AZ                                         */
} lbm_rigid_body;

typedef struct lbm_rigid_state {
  double x[3], q[4], v[3];
  double L[3];            /* angular momentum (world)                             */
  double omega[3];        /* angular velocity at the last half step               */
  double d[3];            /* displacement since attach (not wrapped)              */
  double force[3], torque[3]; /* last measured fluid force and torque             */
  double mass, inertia[3];
} lbm_rigid_state;

/* Attaches a moving body to a dense 3D solve created with body NONE
 * (standalone solves only), classifies the cells at its pose and clears the
 * force history.  Replaces writeBuffer(bodyBuf, packBodyState(...)),
 * main-3d.js:2983-2990. */
int lbm_rigid_attach(lbm_solve* s, const lbm_rigid_body* body);
/* The body's current state (readback of BodyState3D). */
int lbm_rigid_read(lbm_solve* s, lbm_rigid_state* out);

/* ------------------------------------------------------------------------
 * Two-level block-structured AMR: the dense L0 lattice plus one pool of
 * fine tiles (refinement ratio 2).  L0 is cut into blocks of rb^3 cells; a
 * refined block owns a pool slot of fb^3 fine cells, fb = 2 rb + 4 (the
 * interior plus a two-cell ring), d3-amr.mjs:38-75.  Each macro-step runs
 * the manager (every manage_every steps), explode, two fine substeps,
 * coalesce and the covered L0 step, in the order of main-3d.js:3219-3319.
 * Storage is f32 (shifted) on both levels.
 * ------------------------------------------------------------------------ */
typedef struct lbm_amr_desc {
  int32_t abi_version;  /* LBM_ABI_VERSION                                    */
  int32_t q;            /* 19 or 27; 9 for the 2D quadtree (nz = 1, two       */
                        /* levels, explode interface, no body)                */
  int32_t nx, ny, nz;   /* L0 grid, periodic, multiples of rb, at most 1024   */
                        /* per axis (and so is every pool level's parent      */
                        /* grid): L0 coordinates are packed in 10 bits        */
  int32_t rb;           /* L0 cells per block edge (tile interior 2 rb)       */
  int32_t max_slots;    /* pool capacity in tiles                             */
  int32_t collision;    /* LBM_BGK or LBM_TRT                                 */
  float omega;          /* L0 rate 1/tau_c; fine tau_f = 2 tau_c - 1/2        */
  float omega_minus;    /* L0 TRT odd rate; the fine level is BGK             */
  float force[3];       /* Guo body force, applied unchanged on both levels   */
  int32_t manage_every; /* macro-steps between manager passes; 0: static      */
  float refine_at;      /* |curl u| block maximum that refines a block        */
  float keep_at;        /* |curl u| below which a refined block coarsens      */
  int32_t lead;         /* criterion box dilation in L0 cells                 */
  int32_t explode_linear; /* EXPLODE_LINEAR (common_d3_amr_explode.wgsl)      */
  int32_t orphans;      /* ORPHANS (common_d3_amr_coalesce.wgsl)              */
  int32_t interface_interp; /* 1: the interp coarse/fine interface
                         * (?interface=interp, main-3d.js:1266-1277, pass order
                         * :3344-3357): refinement-blind coarse step, ring
                         * re-interpolated from the coarse state at t, ring
                         * cells collide, every tile restricted each step.
                         * 0: explode/coalesce (the default; conservative).  */
  int32_t reflux;       /* ?reflux=1 (main-3d.js:1330-1343): seam mass and
                         * momentum flux correction, common_d3_amr_flux.wgsl +
                         * common_d3_amr_reflux.wgsl; needs interface_interp  */
  float sponge_width;   /* L0 sponge (SPONGE_W, as lbm_desc); <= 0 disables.
                         * The fine step does not relax, so refined blocks
                         * must stay out of the band (upload_topology).       */
  int32_t sponge_axes[3];
  float sponge_u[3];
  lbm_body body;        /* pinned sphere (LBM_BODY_SPHERE) or NONE.  It lives
                         * on the fine level alone (main-3d.js:3127-3136): the
                         * pool step bounces back off it and holds its
                         * interior at feq(1, 0) (common_d3_amr_step1.wgsl:
                         * 234-276, :387-403), and one force pass per
                         * macro-step integrates it with weight dx^2 = 1/4
                         * (common_d3_force_pool.wgsl).  Needs the explode
                         * interface.  With manage_every = 0 an uploaded
                         * topology must cover every L0 cell within 1 of it;
                         * with the manager, every block whose box comes
                         * within 2 L0 cells of the body is kept refined
                         * whatever the field says (the geometry term of
                         * common_d3_manage.wgsl:257-302, OR'ed with the
                         * field at :443-462), and blocks reaching into the
                         * sponge band are never refined by the field.       */
  int32_t explode_dirs; /* directions per explode thread, 1..q; 0: chosen from
                         * the ring size (about 2^20 threads per launch).  A
                         * launch-shape knob: results do not depend on it.    */
  int32_t criterion;    /* the field refine_at / keep_at are thresholds on:
                         * 0 the vorticity magnitude |curl u|; 1 Q = 1/2
                         * (|Omega|^2 - |S|^2) (qAt, common_d3_criterion.wgsl:
                         * 102-120).  Either is the block maximum over the
                         * block's L0 box dilated by `lead`.                 */
  int32_t levels;       /* levels including L0: 2 (0 reads as 2), 3 or 4.
                         * Level m >= 2 refines level m - 1's tiles the way
                         * level 1 refines L0 (pool parents, common_d3_parent_
                         * pool.wgsl; the per-level schedule, main-3d.js:
                         * 3169-3180): tau_m = 2 tau_(m-1) - 1/2, the body on
                         * the finest level alone.  Levels >= 3 take the
                         * explode interface; a static hierarchy is uploaded
                         * level by level, and with manage_every > 0 the
                         * manager decides every level: the body's geometry
                         * at the finest level, the field per level on the
                         * ladder below, then the 2:1 closure (octets and
                         * the 27-neighbour parent gather, common_d3_manage.
                         * wgsl completeOctets/balance), drain finest-first
                         * and fill coarsest-first (main-3d.js:3251-3321). */
  float ladder_octaves; /* the field ladder (desiredLevelAt, common_refine.
                         * wgsl:54-79): each further level costs this many
                         * octaves of rotation rate (N_REFINE_INC, ?qinc=),
                         * so level m wants a block when the criterion
                         * reaches refine_at x r^(m-1) (keep_at x r^(m-1)
                         * for a held one), r = 2^octaves for |curl u| and
                         * 4^octaves for Q (a squared rate).  <= 0 reads as
                         * 1.  The criterion is the L0 field's, per level-1
                         * block, as the reference reads it.                 */
} lbm_amr_desc;

typedef struct lbm_amr lbm_amr;

/* Replaces the per-level pool buffers and pipelines, main-3d.js:775-807,
 * :1223-1700.  Refuses tau_c within 1e-3 of 1 (main-3d.js:430-442). */
int lbm_amr_create(const lbm_amr_desc* desc, lbm_amr** out);
int lbm_amr_destroy(lbm_amr* a);

/* Block maps (blockSlot[nblocks], slotToBlock[max_slots], -1 = none), as
 * refineWhere / refineHierarchy build them (d3-amr.mjs:80-99, :603-651).
 * Replaces uploadLevelTopology, main-3d.js:822. */
int lbm_amr_upload_topology(lbm_amr* a, const int32_t* block_slot, const int32_t* slot_to_block);
int lbm_amr_read_topology(lbm_amr* a, int32_t* block_slot, int32_t* slot_to_block);
/* Level m's block maps (blockSlot over level m - 1's global grid, L0 dims x
 * 2^(m-1) / rb blocks; slotToBlock[max_slots]), as the manager left them
 * (poolStateAt, main-3d.js:4248-4260). */
int lbm_amr_read_level_topology(lbm_amr* a, int32_t level, int32_t* block_slot, int32_t* slot_to_block);

/* Block maps of pool level `level` (1 .. levels - 1): blockSlot over the
 * level's blocks (the parent level's global grid, L0's times 2^(level-1),
 * in blocks of rb), slotToBlock[max_slots].  Level 1 is
 * lbm_amr_upload_topology.  A deeper level's tiles must nest in the parent
 * level's tiles and every ring cell's parent cell must lie in one
 * (checkRingParentCoverage, d3-amr.mjs:348-390); with a body, the finest
 * level must refine every block within one L0 cell of it.  Upload every
 * level before lbm_amr_upload_macro.  Replaces uploadLevelTopology for
 * m >= 2, main-3d.js:822, with refineHierarchy's levels (d3-amr.mjs:603). */
int lbm_amr_upload_level_topology(lbm_amr* a, int32_t level, const int32_t* block_slot,
                                  const int32_t* slot_to_block);

/* L0 [rho, ux, uy, uz] seed -> equilibrium, then every refined tile is
 * filled from L0 by interpolation (reset, main-3d.js:2973-3035), level by
 * level from its parent. */
int lbm_amr_upload_macro(lbm_amr* a, const float* mac4);

/* Advances n macro-steps (encodeSteps, main-3d.js:3122-3335).  Returns
 * LBM_E_NOMEM when the manager could not place a wanted tile (pool
 * exhaustion is a hard failure, common_d3_manage.wgsl:500-525). */
int lbm_amr_step(lbm_amr* a, int32_t n);

/* Runs the manager once now (criterion, decide, drain, coarsen, refine,
 * fill) without stepping. */
int lbm_amr_manage_now(lbm_amr* a);

/* L0 macroscopic field (covered cells hold the coalesced projection of the
 * fine tiles). */
int lbm_amr_read_macro(lbm_amr* a, float* mac4);
/* Pool macroscopic field, planar: [4][max_slots * fb^3]. */
int lbm_amr_read_pool_macro(lbm_amr* a, float* mac_pool);
/* Populations, f32 wire format: L0 [q][ncells]; pool [q][max_slots * fb^3]. */
int lbm_amr_read_f(lbm_amr* a, float* f_l0, float* f_pool);
/* Pool level `level` (1 .. levels - 1): populations, f32 wire format
 * [q][max_slots * fb^3], and the planar macroscopic field [4][max_slots *
 * fb^3] of its last substep.  Interior cells a finer level covers hold that
 * level's outbox and coalesced moments, not a solution of this level. */
int lbm_amr_read_level_f(lbm_amr* a, int32_t level, float* f_pool);
int lbm_amr_read_level_macro(lbm_amr* a, int32_t level, float* mac_pool);
/* Per block |curl u| maximum from the last manager pass, and its decision. */
int lbm_amr_read_criterion(lbm_amr* a, float* crit, uint8_t* want);
/* [count, slot ...] of the slots in use, ascending. */
int lbm_amr_read_active(lbm_amr* a, int32_t* list);
/* Force history of the fine-level body, one FP64 record [fx, fy, fz, tx,
 * ty, tz] per macro-step in L0 units (fine momentum exchange times dx^2 =
 * 1/4, torque about the centre in L0 cells), the most recent max_records,
 * oldest first; as lbm_read_forces.  Replaces the pool force pass's
 * readback (common_d3_force_pool.wgsl, main-3d.js:3138-3141). */
int lbm_amr_read_forces(lbm_amr* a, double* out, int32_t max_records, int32_t* count);
/* FP64 [mass, px, py, pz] of the composite state: unrefined L0 cells plus,
 * at each level m, the interior cells no finer tile covers, weighted 8^-m. */
int lbm_amr_read_totals(lbm_amr* a, double* out4);
/* step count, tiles in use, wanted tiles left unplaced (sticky). */
int lbm_amr_info(lbm_amr* a, int64_t* step, int32_t* active_tiles, int32_t* exhausted);


/* ------------------------------------------------------------------------
 * Batched serving.  No counterpart in the reference (it runs one solve per
 * page); the north-star serving loop.  Many independent dense solves of
 * uneven sizes stay resident in one packed device pool and advance together:
 * every launch of the batched step kernel steps every resident solve of one
 * lattice class (Q, storage) once.  A host worker thread admits queued
 * requests as pool memory frees up (against the byte budget), runs
 * graph-captured chunks of steps, retires finished solves between chunks and
 * copies their results back.  A solve's result does not depend on what else
 * is resident: tiles, rounding and force summation order are its own.
 * ------------------------------------------------------------------------ */
typedef struct lbm_serve_desc {
  int32_t abi_version;   /* LBM_ABI_VERSION                                    */
  int32_t max_resident;  /* resident solves at most (0: limited by the budget)  */
  double budget_bytes;   /* device pool size: admission budget                 */
  int32_t chunk_steps;   /* steps per captured chunk (admission granularity)   */
  int32_t autotune;      /* 1: tile shape per lattice class by timing its first
                            request, when that holds 2^20 cells or more       */
  double target_cells;   /* stop admitting once this many cells are resident
                            (enough to fill the device; 0: no limit)          */
  int32_t persistent;    /* 0: a chunk is a captured graph of chunk_steps
                            batched launches per lattice class; 1: one
                            cooperative persistent launch per class, whose
                            resident blocks step every tile chunk_steps times
                            with a grid barrier between steps              */
} lbm_serve_desc;

typedef struct lbm_serve_result {
  int64_t ticket;
  int32_t status;        /* LBM_OK, or the request's error code                */
  int32_t steps;
  uint32_t cells;
  double t_submit;       /* seconds on the server clock                        */
  double t_admit;
  double t_finish;       /* results copied back                                */
  double force[6];       /* last step's [fx, fy, fz, tx, ty, tz] (FP64)        */
} lbm_serve_result;

typedef struct lbm_serve_stats {
  int64_t submitted, completed, failed, resident, queued;
  double cell_updates;   /* sum of cells x steps over completed solves         */
  double first_submit, last_finish;  /* server clock, seconds                  */
  double p50_latency, p99_latency;   /* submit -> finish over completed, s     */
  double mlups;          /* cell_updates / (last_finish - first_submit) / 1e6  */
  double pool_bytes, peak_resident_bytes;
  int32_t peak_resident;
  int32_t chunks;        /* chunks executed                                    */
  double device_seconds; /* device time of the chunk graphs (HIP events)       */
  double device_bytes;   /* algorithmic bytes they moved: 2 Q s per cell update
                            (s = 4 for f32, 2 for the 16-bit formats)         */
  int64_t graph_instantiations; /* chunk graphs instantiated                    */
  int64_t graph_updates; /* chunk graphs updated in place for a new membership */
} lbm_serve_stats;

typedef struct lbm_serve lbm_serve;

int lbm_serve_create(const lbm_serve_desc* desc, lbm_serve** out);
/* Stops the worker after the requests already running; queued ones fail. */
int lbm_serve_destroy(lbm_serve* sv);
/* Queues a solve: descriptor (autotune ignored), initial [rho, ux, uy, uz]
 * per cell (copied), steps to run.  Refuses (LBM_E_NOMEM) a request larger
 * than the whole budget. */
int lbm_serve_submit(lbm_serve* sv, const lbm_desc* desc, const float* mac4, int32_t steps, int64_t* ticket);
/* Resumable requests: exactly one of mac4 (a fresh start, as above) and
 * state (a state from lbm_save_state or lbm_serve_read_state for the same
 * descriptor; copied) gives the starting point; keep_state 1 keeps the final
 * state for lbm_serve_read_state.  A solve split into requests that hand
 * their states on gives the uninterrupted solve's bits. */
int lbm_serve_submit_state(lbm_serve* sv, const lbm_desc* desc, const float* mac4, const void* state,
                           size_t state_bytes, int32_t steps, int32_t keep_state, int64_t* ticket);
/* The final state of a finished request submitted with keep_state: with buf
 * null only *size is set; else buf must hold *size bytes. */
int lbm_serve_read_state(lbm_serve* sv, int64_t ticket, void* buf, size_t* size);
/* Waits up to timeout_s (< 0: forever) for a request; LBM_E_STATE on timeout. */
int lbm_serve_wait(lbm_serve* sv, int64_t ticket, double timeout_s, lbm_serve_result* out);
/* Results of a finished request: final macroscopic field, force history. */
int lbm_serve_read_macro(lbm_serve* sv, int64_t ticket, float* mac4);
int lbm_serve_read_forces(lbm_serve* sv, int64_t ticket, double* out6, int32_t max_records, int32_t* count);
/* Drops a finished request's host-side results. */
int lbm_serve_release(lbm_serve* sv, int64_t ticket);
/* Waits until every submitted request has finished. */
int lbm_serve_drain(lbm_serve* sv, double timeout_s);
int lbm_serve_read_stats(lbm_serve* sv, lbm_serve_stats* out);
/* Server clock (seconds since create), for callers scheduling arrivals. */
double lbm_serve_clock(lbm_serve* sv);

#ifdef __cplusplus
}
#endif
#endif
