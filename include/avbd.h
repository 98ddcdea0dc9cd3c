/*
 * avbd.h — C ABI of the MI355X AVBD rigid-body step (libavbd_mi355x.so).
 *
 * Drop-in for the reference's WebGPU solver class GpuSolver3D
 * (src/avbd3d/gpu/solver.ts:504), whose operational surface this mirrors. Plain C: pointers,
 * sizes and status codes; no torch or HIP types. One host thread per solver; every call is
 * ordered on the solver's own HIP stream; reads synchronise that stream.
 *
 * Record formats are the reference's (src/avbd3d/gpu/layout.ts) so callers keep their
 * readback code:
 *   body      40 f32  (B_POS 0 .. B_ANGVEL 36; layout.ts:49-58)
 *   joint     40 f32  (J_PEN_LIN 0 .. J_REST_START 36; layout.ts:84-110) + info int32[4]
 *             (type 1 joint / 2 spring / 0 none, body a (-1 world), body b, 0)
 *   manifold  8 u32   (a, b, first contact, count | step << 4; normal xyz, friction as f32 bits)
 *   contact   16 u32  (rA xyz, key | STICK; rB xyz, C0.n; pen xyz, C0.t1; lam xyz, C0.t2)
 * Inside the library bodies live as SoA float4 streams; the 40-float record is the interchange
 * format only.
 *
 * Status codes: 0 ok; negative on error (avbd_last_error gives the message). Invalid input is
 * refused before anything changes (solver.ts:858-875). Capacity overflow during a step is not
 * an error: it is counted in avbd_counters.overflow (bits 1 pairs, 2 contacts, 4 manifolds),
 * as the reference counts it (wgsl-collision.ts:230, :727, :816). Bit AVBD_OVERFLOW_BARRIER is
 * this library's own: see avbd_set_barrier_polls.
 */
#ifndef AVBD_MI355X_H
#define AVBD_MI355X_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AVBD_OK 0
#define AVBD_ERR_INVALID (-1)
#define AVBD_ERR_CAPACITY (-2)
#define AVBD_ERR_DEVICE (-3)
#define AVBD_ERR_DESTROYED (-4)

/* GpuParams3D.flags bits (layout.ts:156-166) */
#define AVBD_FLAG_MATCH_NEAREST 1u
#define AVBD_FLAG_FACE_BIAS 2u
#define AVBD_FLAG_REUSE_CONTACTS 4u
#define AVBD_FLAG_START_AT_REST 8u
#define AVBD_FLAG_MASS_PENALTY 16u
#define AVBD_FLAG_DEFAULTS 31u
/* Host-side: every 10 steps read the counters back and adapt the colour cap and colouring rounds
 * (GpuSolver3D.adaptColors, solver.ts:2379-2407); the read is synchronous, so runs stay
 * deterministic. Not in the reference's flag word (World.step drives its adapt, world.ts:1740). */
#define AVBD_FLAG_ADAPT_COLORS 128u
/* Host-side: fixed storage (the reference's fixed plan, src/lib/storage-budget.ts): capacities
 * stay as created and work past them is dropped and counted. Without it, avbd_adapt (and the
 * readback of AVBD_FLAG_ADAPT_COLORS) grows pairs, manifolds and contacts before they overflow. */
#define AVBD_FLAG_FIXED_STORAGE 256u

#define AVBD_BODY_FLOATS 40
#define AVBD_JOINT_FLOATS 40
#define AVBD_MANIFOLD_WORDS 8
#define AVBD_CONTACT_WORDS 16
#define AVBD_NO_COLOR 255u

typedef struct avbd_solver avbd_solver;

/* Replaces GpuParams3D + the parts of GpuSolverOptions a step reads (solver.ts:175-274). */
typedef struct avbd_params {
  float dt;             /* 1/60 */
  float gravity;        /* along up; negative pulls down */
  float beta_lin;       /* 1e4 */
  float beta_ang;       /* 100 */
  float gamma;          /* 0.999 */
  float alpha;          /* 0.99 */
  float contact_margin; /* 0.01 (COLLISION_MARGIN) */
  float up[3];          /* unit up vector */
  uint32_t flags;       /* AVBD_FLAG_* */
  uint32_t iterations;  /* 10 */
  uint32_t color_cap;   /* colours the colouring may use and the solve sweeps, <= 64 */
  uint32_t color_rounds;/* Jones-Plassmann rounds per step, even */
  uint32_t primal_lanes[4]; /* colour sizes from which 1, 2, 4, 8 threads per body are used (else 16) */
} avbd_params;

/* Replaces GpuSolverOptions.bodyCapacity/capacity (solver.ts:233-274). 0 = default sizing. */
typedef struct avbd_capacity {
  uint32_t bodies, joints, pairs, manifolds, contacts, hull_words;
} avbd_capacity;

/* readCounters (solver.ts:2305-2308). */
typedef struct avbd_counters {
  uint32_t pairs, contacts, manifolds, overflow, clashes, colors, color_cap;
  uint32_t step; /* steps since creation (the contact stamps' clock) */
  uint32_t clashing_colors[2]; /* the last step's colours run serially (bit c of word c / 32) */
} avbd_counters;

/* StepProfile (avbd2d/gpu/layout.ts:289-293) in ms, from hipEvents on the solver's stream. */
typedef struct avbd_profile {
  float collision, adjacency, coloring, solve, total;
} avbd_profile;

/* Per-kernel timing over a profiled run (the roofline's launch durations). */
typedef struct avbd_kernel_stats {
  float primal_ms, dual_ms, narrowphase_ms, broadphase_ms, solve_ms, other_ms;
  uint32_t primal_launches, dual_launches, narrowphase_launches, broadphase_launches, solve_launches;
  uint32_t solve_blocks; /* grid of the cooperative solve kernel (0: per-colour launches) */
  float tail_ms;          /* the fused trailing colours (one workgroup per sweep) */
  uint32_t tail_launches;
} avbd_kernel_stats;

typedef struct avbd_step_opts {
  int reuse_manifolds; /* GpuStepOptions.reuseManifolds (solver.ts:43-50): substeps after the first */
} avbd_step_opts;

/* new GpuSolver3D(device, ref, options) (solver.ts:722): bodies and joints as records. */
int avbd_create(const avbd_params *params, const float *bodies, uint32_t body_count, const float *joints,
                const int32_t *joint_info, uint32_t joint_count, const avbd_capacity *capacity, avbd_solver **out);
/* Batched worlds (no counterpart in the reference, which has one world per solver): `worlds`
 * independent worlds packed back to back, world w holding bodies [world_start[w],
 * world_start[w + 1]). Bodies of different worlds never collide; colouring priorities use each
 * body's index within its world, so every world steps exactly as it would in a solver of its
 * own, while each kernel launch covers all worlds. Joints index bodies globally. */
int avbd_create_batch(const avbd_params *params, const float *bodies, uint32_t body_count, const float *joints,
                      const int32_t *joint_info, uint32_t joint_count, const uint32_t *world_start, uint32_t worlds,
                      const avbd_capacity *capacity, avbd_solver **out);
/* GpuSolver3D.destroy (solver.ts:2439). */
void avbd_destroy(avbd_solver *s);
/* The last error's message (the reference throws Error / StorageCapacityError). */
const char *avbd_last_error(const avbd_solver *s);
/* Message of the last failed avbd_create (no solver exists to carry it). */
const char *avbd_create_error(void);

/* GpuSolver3D.step (solver.ts:2112-2293): one substep, enqueued on the solver's stream. */
int avbd_step(avbd_solver *s, const avbd_step_opts *opts);
/* Several steps; the 1st detects contacts, the rest reuse manifolds when substeps > 1. */
int avbd_steps(avbd_solver *s, uint32_t steps, uint32_t substeps);
/* params (mutable between steps). */
int avbd_set_params(avbd_solver *s, const avbd_params *params);
int avbd_get_params(const avbd_solver *s, avbd_params *params);

/* readBodies / rewriteBodies (solver.ts:2297, :1774): records [first, first+count). */
int avbd_read_bodies(avbd_solver *s, float *out, uint32_t first, uint32_t count);
int avbd_write_bodies(avbd_solver *s, const float *in, uint32_t first, uint32_t count);
/* readBodies without allocating (no counterpart in the reference): for callers that cannot call
 * hipMalloc, such as tight serving or render loops and code that must not allocate. Same records
 * as avbd_read_bodies, staged through a fixed array of AVBD_READ_NOALLOC_MAX body records that the
 * library keeps on each GPU for the life of the process (instead of the solver's lazily grown
 * staging buffer). At most AVBD_READ_NOALLOC_MAX bodies per call (more: AVBD_ERR_INVALID, nothing
 * read); read larger ranges in several calls. Every solver on one GPU shares that array: calls
 * on the same GPU that overlap (solvers driven from different host threads) race. */
#define AVBD_READ_NOALLOC_MAX 16384u
int avbd_read_bodies_noalloc(avbd_solver *s, float *out, uint32_t first, uint32_t count);
/* readJoints (solver.ts:2300) */
int avbd_read_joints(avbd_solver *s, float *out, uint32_t first, uint32_t count);
/* readCounters (solver.ts:2305) */
int avbd_read_counters(avbd_solver *s, avbd_counters *out);
/* readContactList (solver.ts:2311): manifolds then contacts of the last step, raw records. */
int avbd_read_contacts(avbd_solver *s, uint32_t *manifolds, uint32_t manifold_cap, uint32_t *contacts, uint32_t contact_cap);
/* readPairs (solver.ts:2348): broadphase pairs (a > b) of the last detecting step. */
int avbd_read_pairs(avbd_solver *s, uint32_t *out, uint32_t cap);
/* Colour state per body (colour | PENDING<<8) and the colour buckets' starts (65 words). */
int avbd_read_colors(avbd_solver *s, uint32_t *state, uint32_t count, uint32_t *starts);

/* The reference poses updateRefs keeps per body (wgsl-collision.ts:34-48; its `refs` buffer), 8
 * floats each: pos xyz, 0, rot xyzw (all zero: none yet). A validation read: with the bodies,
 * contacts, colour state and step count it is the whole state the next step depends on. */
int avbd_read_refs(avbd_solver *s, float *out, uint32_t first, uint32_t count);
/* seedFrom (solver.ts:1320): bodies, joints with warm-start state, and last step's contacts. */
int avbd_seed(avbd_solver *s, const float *bodies, uint32_t body_count, const float *joints, uint32_t joint_count,
              const uint32_t *manifolds, uint32_t manifold_count, const uint32_t *contacts, uint32_t contact_count);
/* fixedColors (solver.ts:560): colours per body, NO_COLOR for none; NULL returns to colouring. */
int avbd_set_fixed_colors(avbd_solver *s, const uint32_t *colors, uint32_t count);
/* Colour state the next colouring starts from (validation of the colouring on given state). */
int avbd_set_color_state(avbd_solver *s, const uint32_t *state, uint32_t count);
/* IgnoreCollision pairs (forces.ts:384; NoCollideTable.add with IGNORE_SLOT). */
int avbd_add_ignore_pairs(avbd_solver *s, const uint32_t *pairs, uint32_t count);
/* setFilters (solver.ts:1791): collision groups and the groups each collides with. */
int avbd_set_filters(avbd_solver *s, const uint32_t *groups, const uint32_t *collides_with, uint32_t count);
/* Hull buffer (wgsl-hull.ts header format, vec4u words); bodies reference it by shape code. */
int avbd_set_hulls(avbd_solver *s, const uint32_t *words, uint32_t word_count);

/* adapt (solver.ts:2362-2407). Colouring half: colour cap = colours in use + 3 spare, grown at
 * once on clashes or a spent spare, shrunk after three quiet reads; rounds 4..32. Capacity half
 * (not with AVBD_FLAG_FIXED_STORAGE or fixed colours, nor inside avbd_forward/avbd_backward): a
 * buffer whose count overflowed or passed 0.8 of its capacity grows to max(2 x count, 1.25 x
 * capacity), keeping the warm-start contacts; the adjoint's buffers and recorded rollout are
 * dropped. */
int avbd_adapt(avbd_solver *s, const avbd_counters *counters);
/* contactCapacity / pairCapacity / manifoldCapacity (solver.ts) and the fixed ones. */
int avbd_get_capacity(const avbd_solver *s, avbd_capacity *out);
/* The broadphase configuration uploadStatics chose (solver.ts:1263-1293): cell size, the fine
 * level's cell size and radius threshold (equal to the cell size and the large-body threshold when
 * there is no fine level), whether the coarse level holds moving bodies, and the large bodies. */
int avbd_broadphase_info(const avbd_solver *s, float *cell_size, float *fine_cell_size, float *fine_radius, uint32_t *coarse_movers,
                         uint32_t *large_count);

/* How the solve pass is scheduled (results are identical whichever runs). */
typedef struct avbd_schedule {
  uint32_t cooperative;      /* 0: per-colour launches; 1: one cooperative launch per substep with grid
                                barriers; 2: the dataflow sweep, one cooperative launch per substep whose
                                bodies wait only for their neighbours' per-body completion counts;
                                3: per-colour launches, then the dataflow tail (see avbd_set_sparse_tail);
                                4: as 3 with the colours before the tail as one dataflow launch per
                                iteration (a workgroup a chunk, in colour order) */
  uint32_t threads;          /* workgroup size of the cooperative kernel: 64..1024 (the dataflow sweep: 64) */
  uint32_t blocks_per_cu;    /* workgroups per CU of the cooperative kernel */
} avbd_schedule;
int avbd_set_schedule(avbd_solver *s, const avbd_schedule *schedule);
int avbd_get_schedule(const avbd_solver *s, avbd_schedule *schedule);
/* The cooperative solve's grid barrier waits a bounded number of polls (default 2^22, ~0.1 s),
 * then gives up rather than hang the device (no counterpart in the reference, which has no grid
 * barrier). A give-up means that launch's solve ran partly unordered, so its step's result is not
 * the step's: the solver then switches to per-colour launches for good, the avbd_step/avbd_steps
 * call whose readback saw it (with AVBD_FLAG_ADAPT_COLORS, every 10 steps) returns
 * AVBD_ERR_DEVICE, and the next avbd_read_counters sets AVBD_OVERFLOW_BARRIER in overflow. An
 * adjoint rollout that saw one refuses its backward. `polls` 0 restores the default; a tiny value
 * forces the give-up path (tests). */
#define AVBD_OVERFLOW_BARRIER 256u
int avbd_set_barrier_polls(avbd_solver *s, uint32_t polls);
/* fuseSparseColors / primalSparse (solver.ts:574, :2133-2143, :2248-2255; wgsl-solve.ts:378-394):
 * with per-colour launches, colours from `first_colour` up to the cap run in one workgroup, in
 * order, instead of a launch each. The reference starts that tail at colour 16 for scenes of at
 * most ~1k bodies. Here the start is chosen: a colour index, 64 for none, or AVBD_TAIL_AUTO
 * (the default), which picks the first of the trailing colours that held <= 32 bodies at each
 * adaptive-colouring readback (AVBD_FLAG_ADAPT_COLORS; without that flag, no tail). Under schedule
 * 3 and 4 the same start opens the dataflow tail (the colours from it and the dual in one launch
 * per iteration); its automatic start is the first of the trailing colours that held <= 2048
 * bodies. Results are the same whichever runs. */
#define AVBD_TAIL_AUTO 0xFFFFFFFFu
int avbd_set_sparse_tail(avbd_solver *s, uint32_t first_colour);

/* World's command passes on body records between steps; one record per body per call (a body
 * commanded twice is refused, as the reference's host sums a body's pushes into one record).
 * Pusher (src/lib/push.ts:28-45, apply :90-140): 12 floats per pushed body: the linear impulse xyz
 * and the body index (u32 bits), the part applied off-centre xyz and 0, its moment about the origin
 * plus angular impulses xyz and 0. velocity += lin / m, angular velocity += I^-1 (moment - x x
 * offLin) in the principal frame; static and disabled bodies are left alone. */
int avbd_push(avbd_solver *s, const float *pushes, uint32_t count);
/* BodyMover (src/avbd3d/gpu/body-mover.ts:10-31, apply :56-74): 12 floats per moved body: target
 * position xyz and the body index (u32 bits), target rotation xyzw, steps left, dt, 1 when the
 * rotation is given (else 0), 0. Sets the velocity that reaches the target in `left` steps and the
 * angular velocity whose step covers 1/left of the shortest arc (zero without a rotation). */
int avbd_move(avbd_solver *s, const float *moves, uint32_t count);
/* BodyReset (src/avbd3d/gpu/body-reset.ts:9-17, apply :41-52): x- and the inertial pose to the
 * body's current pose, velocity and angular velocity to zero. */
int avbd_reset(avbd_solver *s, const uint32_t *indices, uint32_t count);

/* Contact events, ContactWatch (src/lib/contacts.ts:134-270; World drives it at world.ts:1736-1737):
 * after a step, its manifolds whose bodies both carry a lifetime ID and one of which reports
 * contacts go into this step's hash set of ID pairs. A pair missing from the last run's set
 * began touching; one of the last run's set missing from this one ended. */
typedef struct avbd_contact_event {
  uint32_t kind;      /* AVBD_EVENT_BEGIN or AVBD_EVENT_END */
  uint32_t a, b;      /* lifetime IDs; a belongs to the pair's higher solver slot */
  uint32_t step;      /* the step passed to avbd_watch_run */
  float point[3];     /* begin: mean of the contact points on body a, world frame */
  float impulse;      /* begin: |sum of the points' normal lambdas| x dt */
  float normal[3];    /* begin: the manifold normal (from b to a) */
  float pad;
} avbd_contact_event;
#define AVBD_EVENT_BEGIN 1u
#define AVBD_EVENT_END 2u
/* The constructor (contacts.ts:158-177): two sets of 2^ceil(log2(max(64, 2 pairs))) slots and room
 * for `capacity` events between reads. A later call replaces the sets and events (the next run
 * reports every touching pair as begun, as the reference's growContactWatch, world.ts:1779-1790)
 * and keeps the watched bodies. */
int avbd_watch_create(avbd_solver *s, uint32_t pairs, uint32_t capacity);
/* setWatched (contacts.ts:193-196): body indices[k] gets lifetime ID ids[k] (0: none) and reports
 * its contacts when on[k] is nonzero. */
int avbd_watch_set(avbd_solver *s, const uint32_t *indices, const uint32_t *ids, const uint8_t *on, uint32_t count);
/* run (contacts.ts:199-217): the last step's first `pairs` manifolds against the last run's set. */
int avbd_watch_run(avbd_solver *s, uint32_t pairs, uint32_t step, float dt);
/* read (contacts.ts:224-261): the events since the last read in the order written (at most
 * `capacity` copied to out), how many did not fit (the watch's capacity, then out's), and whether
 * a run since the last read had more touching pairs than a set holds. Clears them. */
int avbd_watch_read(avbd_solver *s, avbd_contact_event *out, uint32_t capacity, uint32_t *count, uint32_t *dropped, uint32_t *full);
/* ---- Scene edits: what World.flush (src/lib/world.ts:1861-1995) calls between steps ----------
 * Each is refused before anything changes when its input is invalid (AVBD_ERR_INVALID) or, with
 * AVBD_FLAG_FIXED_STORAGE, does not fit (AVBD_ERR_CAPACITY, the reference's StorageCapacityError).
 * An edit makes the next step detect contacts afresh (the reference's manifoldsValid = false). */
/* jointInfo (solver.ts:2356): per joint slot its type (1 joint, 2 spring, 0 none/released), body a
 * (-1: the world), body b, 0. */
int avbd_read_joint_info(const avbd_solver *s, int32_t *out, uint32_t first, uint32_t count);
/* Body and joint counts now (bodyCount, jointCount: slots in use or released). */
int avbd_get_sizes(const avbd_solver *s, uint32_t *bodies, uint32_t *joints);
/* addBodies (solver.ts:1544): appends body records after the last body, as many as fit in the body
 * capacity (avbd_capacity.bodies); returns how many were added and sets *first to the first one's
 * index (-1 when none fit). New bodies start uncoloured; a record's x- .w word (1: a new occupant,
 * no contact history) is its lifecycle word, as the reference's pack writes it. Not on a batched
 * solver (its worlds are fixed). */
int avbd_add_bodies(avbd_solver *s, const float *bodies, uint32_t count, int32_t *first);
/* patchBody (solver.ts:1103): from `record` (a whole 40-float body record), the fields in `fields`
 * (AVBD_FIELD_*): position writes pos, x- and the inertial position; rotation rot, q- and the
 * inertial rotation; velocity and angular velocity their xyz. `fixed`: also mass and moments (a
 * body made fixed or dynamic; it starts uncoloured). `marker`: the lifecycle word becomes
 * "rewritten" with the body's current lifecycle bits. */
#define AVBD_FIELD_POSITION 1u
#define AVBD_FIELD_ROTATION 2u
#define AVBD_FIELD_VELOCITY 4u
#define AVBD_FIELD_ANGULAR_VELOCITY 8u
int avbd_patch_body(avbd_solver *s, uint32_t index, uint32_t fields, const float *record, int fixed, int marker);
/* rewriteFixed (solver.ts:1140): bodies indices[k] become fixed boxes of `size` (3 floats, full
 * widths) and `friction` at positions[3k..] turned by rotations[4k..] (NULL: unturned), at rest. */
int avbd_rewrite_fixed(avbd_solver *s, const uint32_t *indices, uint32_t count, const float *positions, const float *size,
                       float friction, const float *rotations);
/* JointSpec (solver.ts:60-124). NaN marks a field left out. */
typedef struct avbd_joint_spec {
  int32_t a, b;        /* body indices */
  float rA[3], rB[3];  /* anchors, body-local */
  float angular;       /* angle-lock stiffness (0: a ball joint); NaN: rigid */
  float fracture;      /* NaN: the call's; INFINITY: never */
  int32_t linear;      /* breaks on the linear force too: 0/1, -1: the call's */
  float rest[4];       /* the turn the angle lock holds (x, y, z, w, normalised); all 0: none */
  float yield;         /* plastic: angular force past which it bends and keeps the bend; NaN: none */
  float break_bend;    /* plastic: bend (rad, 0..pi) past which it breaks; NaN/INFINITY: never */
} avbd_joint_spec;
/* appendJoints (solver.ts:1637): joints into released slots (lowest first) then new ones, written
 * in one upload; slots[k] receives spec k's slot. The joint storage grows unless fixed. */
int avbd_append_joints(avbd_solver *s, const avbd_joint_spec *specs, uint32_t count, float fracture, int linear, uint32_t *slots);
/* appendJoint (solver.ts:1559): one joint with the given stiffnesses (INFINITY: rigid), a = -1 for a
 * world point rA; appended after the last slot unless the storage is fixed. */
int avbd_append_joint(avbd_solver *s, int32_t a, int32_t b, const float *rA, const float *rB, float stiff_lin, float stiff_ang,
                      uint32_t *slot);
/* appendWorldJoint (solver.ts:1592), its 'legacy' policy: a hard ball joint from a world point to a
 * point of body b (the mouse drag). */
int avbd_append_world_joint(avbd_solver *s, uint32_t b, const float *anchor_world, const float *anchor_body, uint32_t *slot);
/* appendSprings (solver.ts:1685): springs pulling rA on a towards rB on b (a = -1: a world point)
 * with `stiffness` (N/m) about `rest` metres apart. */
typedef struct avbd_spring_spec {
  int32_t a, b;
  float rA[3], rB[3];
  float stiffness, rest;
} avbd_spring_spec;
int avbd_append_springs(avbd_solver *s, const avbd_spring_spec *springs, uint32_t count, uint32_t *slots);
/* releaseJoints (solver.ts:1745): slots switched off and zeroed, their bodies colliding again, the
 * slots reused by later appends; released or unknown slots are ignored. */
int avbd_release_joints(avbd_solver *s, const uint32_t *slots, uint32_t count);
/* setWorldAnchor (solver.ts:1937): a world joint's world point (3 floats). */
int avbd_set_world_anchor(avbd_solver *s, uint32_t slot, const float *p);
/* disableConstraint (solver.ts:1942): penalties, stiffnesses and multipliers zeroed (fracture and
 * torque arm kept): the joint stops acting. */
int avbd_disable_constraint(avbd_solver *s, uint32_t slot);
/* setBodyAux (solver.ts:1814): body indices[k] falls with gravity x scales[k] and, if dynamic, its
 * velocities decay by exp(-rate dt) per step (linear[k], angular[k] in 1/s; both NULL: a run of
 * more than one slot resets the rates to 0, a single slot keeps them), from the next step. */
int avbd_set_body_aux(avbd_solver *s, const uint32_t *indices, uint32_t count, const float *scales, const float *linear,
                      const float *angular);
/* setLifecycle (solver.ts:2069): lifecycle bits 0, 2 (disabled), 4 (cold) or 6; the body starts
 * uncoloured and its contact history goes. */
int avbd_set_lifecycle(avbd_solver *s, uint32_t index, uint32_t bits);

/* Project autotuner: times each candidate schedule over `steps` x `substeps` (the simulation
 * advances), twice interleaved, and keeps the fastest; writes it and its ms per step to `best`.
 * When a schedule with per-colour launches wins (all but the cooperative sweep), it also times,
 * against the params' primal_lanes rule, eight threads per body for every colour above the rule's
 * last size, each of the two without sixteen-thread colours, and eight threads with sixteen up to
 * four times the last size, and keeps the fastest in the params (avbd_get_params reads it). */
int avbd_autotune(avbd_solver *s, uint32_t steps, uint32_t substeps, avbd_schedule *best, float *best_ms);

/* profileNextStep (solver.ts:1952): phase times of the next step, synchronously. */
int avbd_profile_step(avbd_solver *s, const avbd_step_opts *opts, avbd_profile *out);
/* Runs `steps` x `substeps` with hipEvents around every launch of the hot kernels. */
int avbd_profile_kernels(avbd_solver *s, uint32_t steps, uint32_t substeps, avbd_kernel_stats *out);
/* Times `steps` x `substeps` between two hipEvents on the solver's stream (no readback inside). */
int avbd_time_steps(avbd_solver *s, uint32_t steps, uint32_t substeps, float *ms);
/* Waits for the solver's stream. */
int avbd_sync(avbd_solver *s);

/* Adjoint (no counterpart in the reference, which has no differentiation): the step with its
 * active set frozen (contact set, anchors, normals, warm-start sources, colouring, per-iteration
 * penalties, clamp/cone and warm-start branches, pivot floor), differentiated in f64 at the f32
 * forward's own states. Contact-only scenes (joints and springs are refused).
 * avbd_forward runs `steps` detecting steps (as avbd_step) keeping a checkpoint every
 * `checkpoint_every` steps (0: record every step's tape during the rollout, no re-run later);
 * a capacity overflow or a clashing colouring during it makes the next avbd_backward fail. */
#define AVBD_ADJ_FLOATS 16 /* pos xyz, previous up-velocity, rot xyzw, vel xyz, 0, angvel xyz, 0 */
int avbd_forward(avbd_solver *s, uint32_t steps, uint32_t checkpoint_every);
/* Reverse sweep of the last avbd_forward (re-running each segment from its checkpoint): from the
 * adjoint of the final state (AVBD_ADJ_FLOATS doubles per body) to the adjoint of the initial
 * state (same layout) and of each body's friction (NULL to skip). The solver ends where the
 * rollout ended. */
int avbd_backward(avbd_solver *s, const double *d_final, double *d_initial, double *d_friction);

/* Per-substep hipGraphs (no counterpart in the reference, whose WebGPU command buffers are encoded
 * per step): with `on`, each substep of avbd_step/avbd_steps runs as one graph launch. A substep is
 * captured the first time its launch sequence occurs (the host's every input to it: buffers and
 * capacities, params and colour cap, schedule, sparse tail, reuse) and replayed whenever that
 * sequence recurs; a few such graphs are kept per solver. Results equal the plain launches' bit
 * for bit (the step clock the contact stamps use lives on the device). Rollouts recorded for the
 * adjoint and profiled steps launch plainly, and so does a cooperative solve the runtime will not
 * capture. `on` = 0 drops the graphs. */
int avbd_set_graphs(avbd_solver *s, int on);
/* Substeps captured into graphs and substeps replayed from them since creation. */
int avbd_graph_stats(const avbd_solver *s, uint64_t *captures, uint64_t *replays);

/* Scene planning before a solver exists, as GpuSolver3D's constructor does it from the reference
 * scene (host code: no GPU is touched). */
/* mortonOrder (solver.ts:463-485): order[g] = the index into `bodies` of the body placed g-th
 * along a Z-order curve of the positions (10 bits per axis over their bounding box), ties by
 * index; the reference writes the scene in this order unless spatialSort is false (:800). */
int avbd_morton_order(const float *bodies, uint32_t body_count, uint32_t *order);
/* estimatePairs (solver.ts:350-460): the broadphase's pairs for these bodies (bounding spheres and
 * world AABBs overlap, not both static) counted on a CPU hash grid, how many of them touch (the
 * separating-axis test, 1e-3 tolerance), and each body's touching pairs added into `degree`
 * (body_count ints, may be NULL). */
int avbd_estimate_pairs(const float *bodies, uint32_t body_count, uint64_t *pairs, uint64_t *touching, int32_t *degree);
/* The constructor's sizing from that estimate (solver.ts:745, :815-843): body capacity
 * (least->bodies, else bodies + 1024), joints (+256 spare), pairs max(4096, 4 cap, 2 pairs),
 * manifolds max(4096, cap, 1.6 touching) but at most the pairs, contacts max(8192, 4 cap,
 * 8 touching), each at least least's (NULL: none); and the starting colour cap
 * min(64, max(8, max degree of a dynamic body over touching pairs and joints + 4, least_colors)).
 * joint_info: (type, a, b, 0) per joint, a = -1 for a world joint. */
int avbd_plan_capacity(const float *bodies, uint32_t body_count, const int32_t *joint_info, uint32_t joint_count,
                       const avbd_capacity *least, uint32_t least_colors, avbd_capacity *out, uint32_t *color_cap);

/* ---- SDF collision backend: SdfCollisionBridge (src/avbd3d/gpu/sdf-bridge.ts) ---------------
 * The reference's other producer of the manifold/contact ABI. Installed with
 * avbd_set_collision_backend (setCollisionBackend, solver.ts:2078), it replaces ALL collision
 * detection: each step prepares every body's signed distance field at the body's pose, takes the
 * candidate pairs (given by the caller, or generated on the GPU by a spatial grid), runs Algorithm
 * 1 per pair (cached points advected and re-optimised, fresh samples from the field seeds, first-wins
 * dedup within the tolerance, stable rank order) and exports one manifold per active point into the
 * solver's contact storage, warm-started from the solved contact the point's previous self was
 * exported to (its lambda rotated into the new basis, its penalties, and its stick anchors unless
 * the surface rolls). Solver masses, inertia and poses stay authoritative: a field only gives
 * collision geometry. Not with fixed storage, batched worlds or the adjoint. */
typedef struct avbd_sdf avbd_sdf;
/* Field kinds (packSdfAtlas kinds, sdf-layout.ts:29). */
#define AVBD_SDF_BOX 0u         /* params: half extents xyz */
#define AVBD_SDF_SPHERE 1u      /* radius */
#define AVBD_SDF_CAPSULE 2u     /* radius, half length (along y) */
#define AVBD_SDF_CYLINDER 3u    /* radius, half height (along y) */
#define AVBD_SDF_CONE 4u        /* radius, half height */
#define AVBD_SDF_TORUS 5u       /* major radius, minor radius (about y) */
#define AVBD_SDF_WARP 6u        /* half extents xyz, k; params[4] = type (0 bend, 1 twist, 2 displace) */
#define AVBD_SDF_VOXEL 7u       /* dims, origin, cell, grid (4 floats per node: phi, gradient xyz; x fastest) */
#define AVBD_SDF_SPIKE_PLATE 8u /* half extents xyz, count; spike radius, min height, max height, 0 */
#define AVBD_SDF_SPIKY_WHEEL 9u /* hub radius, hub half width, spike radius, spike half height; spikes */
/* One field (the reference's Field: AnalyticField / VoxelField, ref/sdf/shapes.ts, voxel.ts), in
 * its local frame: bounds, centre of mass (field-local; a body's position is its centre of mass),
 * the kind's parameters (the atlas header words 20..27), the curvature seeds (finalizeSdf's,
 * field-local xyz) and, for meshVertexContacts, the mesh's unique surface vertices. */
typedef struct avbd_sdf_field {
  uint32_t kind;
  uint32_t convex;          /* 1: convex (box, sphere, capsule, cylinder, cone) */
  float bounds_min[3], bounds_max[3];
  float center_of_mass[3];
  float params[8];
  uint32_t dims[3];         /* voxel only */
  float origin[3];
  float cell;
  const float *grid;        /* voxel: 4 * dims[0] * dims[1] * dims[2] floats */
  const float *seeds;       /* 3 per seed */
  uint32_t seed_count;
  const float *supports;    /* 3 per vertex */
  uint32_t support_count;
} avbd_sdf_field;
/* CollisionParams (ref/sdf/manifold.ts:13-28); defaults from avbd_sdf_default_params. */
typedef struct avbd_sdf_params {
  uint32_t optimizer;          /* 0 Nelder-Mead, 1 deep-cut ellipsoid, 2 stochastic gradient descent */
  uint32_t cache_size;         /* cached points re-optimised per pair (4; <= points - 1) */
  uint32_t random_samples;     /* fresh samples per pair and step (1; cache + samples <= points) */
  uint32_t warm_start;         /* 1: keep points across steps */
  uint32_t curvature_sampling; /* 1: fresh samples prefer the fields' curvature seeds */
  uint32_t skip_convex_samples;/* 1: no fresh sample when a convex pair's cache is full and active */
  float epsilon;               /* softmax smoothing (0.1, a length squared) */
  float alpha;                 /* repulsion between points (0.01) */
  float replacement_threshold; /* rank bonus of a cached point (0.01) */
  float tolerance;             /* optimiser and dedup tolerance (0.01) */
  uint32_t max_iterations;     /* 50, 1..512 */
  float step_size;             /* gradient descent only (0.1) */
  float margin;                /* activation distance (0.02) */
  float allowed_penetration;   /* resting gap (0.01) */
  uint32_t signed_separation;  /* signed projection against static boxes */
  uint32_t mesh_vertex_contacts;    /* exact vertex witnesses (fields with supports) */
  uint32_t refresh_deeper_contacts; /* drop a sticking anchor its relocated point is deeper than */
  uint32_t sgd_gradient_stop_only;  /* gradient descent stops on the gradient or the limit only */
} avbd_sdf_params;
void avbd_sdf_default_params(avbd_sdf_params *out);
/* SdfCollisionBridge.create / createAutomatic (sdf-bridge.ts:48-82) for the solver's bodies, which
 * must all exist (one field per body slot: body_field[i] indexes `fields`; body_scale NULL: 1;
 * body_rotation, 4 floats per body xyzw, NULL: none, turns the field in the body frame).
 * automatic = 0: explicit candidates (avbd_sdf_set_pairs), at most max_pairs (0: 1). automatic = 1:
 * the GPU spatial grid generates them each step, at most max_pairs (0: min(n(n-1)/2, max(64, 8n)))
 * with cells of cell_size (0: the median field bounds diagonal); past the capacity they are dropped
 * and counted (overflow bit 1). points: points per pair result, 2..20 (0: 5). The solver's contact
 * and manifold storage is grown to max_pairs x points (automatic), or must hold it (explicit). */
int avbd_sdf_create(avbd_solver *s, const avbd_sdf_field *fields, uint32_t field_count, const uint32_t *body_field,
                    const float *body_scale, const float *body_rotation, const avbd_sdf_params *params, int automatic,
                    uint32_t max_pairs, uint32_t points, float cell_size, avbd_sdf **out);
/* setPairs (sdf-bridge.ts:133): explicit candidates, 3 words each (a, b, key) with a > b, unique
 * pairs and keys; a changed key starts the pair's history afresh. */
int avbd_sdf_set_pairs(avbd_sdf *g, const uint32_t *pairs, uint32_t count);
/* bridge.params (mutable between steps; validated at the next step, as prepare does). */
int avbd_sdf_set_params(avbd_sdf *g, const avbd_sdf_params *params);
/* setCollisionBackend (solver.ts:2078-2091): installs the bridge (NULL: the built-in collision
 * again). The solver then owns it: a replaced bridge and the solver's own end destroy it. The next
 * step detects afresh; a solver with fixed storage, batched worlds or pending lifecycle bits
 * refuses. Graph replays are off while a bridge is installed. */
int avbd_set_collision_backend(avbd_solver *s, avbd_sdf *g);
/* readResults (sdf-bridge.ts:216): the last step's pair results, (8 + 28 points) words each:
 * a, b, key, point count, queries, iterations, id counter, rolling; then per point x xyz, depth,
 * normal xyz, separation, anchor A xyz, softmax, anchor B xyz, rank, active, cached, id, exported
 * contact + 1, ellipsoid e00 e01 e02 e11 e12 e22 0 0. Slots of no pair have a = 0xffffffff. *count
 * receives the slots (explicit: the pairs; automatic: the capacity). */
int avbd_sdf_read_results(avbd_sdf *g, uint32_t *out, uint32_t cap, uint32_t *count);
/* readBroadphaseStats (sdf-bridge.ts:217): pairs found, candidates tested, large bodies, the pair
 * capacity, cell size (automatic only). */
int avbd_sdf_read_stats(avbd_sdf *g, uint32_t *pairs, uint32_t *tested, uint32_t *large, uint32_t *capacity, float *cell_size);
/* grown (sdf-bridge.ts:183-201), in place: the automatic pair capacity raised to max(required,
 * 2 x capacity), the pair results kept (history survives), the solver's contact storage grown. */
int avbd_sdf_grow(avbd_sdf *g, uint32_t required);
/* destroy (sdf-bridge.ts:219); an installed bridge is detached first. A bridge lives no longer
 * than its solver: avbd_destroy frees every bridge made for it (do not destroy those again). */
void avbd_sdf_destroy(avbd_sdf *g);
const char *avbd_sdf_last_error(const avbd_sdf *g);

#ifdef __cplusplus
}
#endif
#endif
