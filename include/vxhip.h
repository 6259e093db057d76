/*
 * vxhip.h -- C ABI of libvxhip, the MI355X (gfx950) device step for Velox.
 *
 * The library plays the role of the reference's device backend
 * (src/backend_cuda.cu behind velox::Backend, include/velox/backend.h:111-172)
 * plus the host-side stages of World::stepImpl that surround it
 * (src/world.cpp:3830-4397).  Records cross the boundary in Velox's own byte
 * layout (include/vx_records.h), so a velox::Backend implementation can pass
 * `std::vector<Body>::data()` and `std::vector<Contact>::data()` straight in.
 *
 * Conventions
 *   - Every entry point returns a status code: VXHIP_OK, VXHIP_ERR_MEMORY
 *     (maps to BackendFailure{MemoryAllocation}), VXHIP_ERR_DEVICE_LOST (maps to
 *     BackendFailure{DeviceLost}), VXHIP_ERR_INVALID (bad argument or scene),
 *     VXHIP_ERR_CAPACITY (VeloxCapacityExceeded), VXHIP_ERR_UNAVAILABLE (no
 *     gfx950 device; the reference factory's nullptr, backend.h:174-176).
 *     The classification follows backend_cuda.cu:24-40.  No C++ exception
 *     crosses the ABI.  vxhip_last_error() describes the latest failure on
 *     the calling thread.
 *   - Host buffers are owned by the caller; device buffers by the context.
 *   - A context is used from one thread at a time (velox_c.h:20-22).
 *   - Setting VELOX_HIP_TEST_FAIL=oom or =device-loss makes the next stage
 *     entry point fail with the matching code without touching the device
 *     (the host-forced fault hook of backend_cuda.cu:69-78).
 */
#ifndef VXHIP_H
#define VXHIP_H

#include "vx_records.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VXHIP_OK = 0,
    VXHIP_ERR_MEMORY = 1,
    VXHIP_ERR_DEVICE_LOST = 2,
    VXHIP_ERR_INVALID = 3,
    VXHIP_ERR_CAPACITY = 4,
    VXHIP_ERR_UNAVAILABLE = 5
};

/* BackendType value this library adds to velox::BackendType and
 * velox_BackendType (appended after Vulkan = 3; backend.h:85-91, velox_c.h:50-55). */
#define VXHIP_BACKEND_TYPE 4

typedef struct vxhip_ctx vxhip_ctx;

typedef struct {
    float gravity[3];          /* World::gravity (world.h:434) */
    int32_t substeps;          /* World::substeps (world.h:438) */
    vx_solver_options options; /* World::solverOptions() (solver.h:23-35) */
} vxhip_config;

typedef struct {
    uint32_t bodies;
    uint32_t pairs;            /* broad-phase candidate pairs */
    uint32_t contacts;         /* narrow-phase contacts */
    uint32_t colors;           /* greedy colors */
    uint32_t ccdCandidates;    /* pairs examined by the conservative-advancement net */
    uint32_t ccdRescued;       /* pairs rewound to their time of impact */
    uint32_t velocityIterations; /* sweeps run in the last step (lastVelocityIterations) */
    uint32_t oversize;         /* bodies on the brute-force broad-phase path */
    float cellSize;            /* broad-phase grid cell edge */
    uint32_t islands;          /* small contact islands solved whole by one workgroup */
    uint32_t islandContacts;   /* contacts in those islands */
    uint32_t bigIslandContacts;/* contacts solved colour by colour (islands above the size limit) */
    uint32_t largestIsland;    /* contacts in the largest island */
    uint32_t maxBatchPasses;   /* most solver passes (colour runs x chunks) one batch needs per sweep */
    uint32_t multiToiEvents;   /* StepStats::multiToiEvents (processHighToi events) */
    uint32_t ccdFlagged;       /* pairs whose end-of-step gap the net flagged */
    uint32_t ccdRelevant;      /* pairs the ordered net walked (flagged and what can follow them) */
    uint32_t warmTies;         /* contacts whose warm-start match a tie among a pair's cached contacts
                                  decided (the cache was then put in the reference's std::sort order) */
    uint32_t launches[6];      /* launches of the last step by kernel, indexed by VXHIP_LAUNCH_* */
    uint32_t graphCaptures;    /* substep graphs captured since the context was created */
    uint32_t graphReplays;     /* substep graph launches beyond the capturing one, since then */
} vxhip_stats;

/* Kernel launch counts in vxhip_stats.launches (one step). */
enum {
    VXHIP_LAUNCH_BATCH = 0,         /* kBatchSweeps: island batches in LDS (warm start and sweeps) */
    VXHIP_LAUNCH_ISLAND = 1,        /* kIslandSweeps: one island of up to the island limit per workgroup */
    VXHIP_LAUNCH_COLOR_PAIR = 2,    /* kSweepColorPair: one big-island colour over the device */
    VXHIP_LAUNCH_COLOR_RUN = 3,     /* kSweepColorRun: a run of narrow colours in one workgroup */
    VXHIP_LAUNCH_TAIL = 4,          /* kSweepTailPair: the colour tail with LDS-resident velocities */
    VXHIP_LAUNCH_GREEDY_DEVICE = 5, /* kGreedyColorDevice: the colouring dataflow across the device */
    VXHIP_LAUNCH_KINDS = 6
};

/* ---------------------------------------------------------------- setup -- */

/* Number of usable gfx950 devices.  Replaces createCudaBackend()'s device probe
 * (backend_cuda.cu:352-356). */
int vxhip_device_count(int* count);

/* Last failure text for the calling thread (never NULL). */
const char* vxhip_last_error(void);

/* Creates a context on the current device.  Replaces createCudaBackend()
 * (backend.h:175). */
int vxhip_create(vxhip_ctx** out);

/* Releases every device buffer; never fails (CudaBackend::~CudaBackend,
 * backend_cuda.cu:358-388). */
void vxhip_destroy(vxhip_ctx* ctx);

/* ------------------------------------------- velox::Backend equivalents -- */

/* Backend::integrate (backend.h:115): velocity integration of `bodies` in place. */
int vxhip_integrate(vxhip_ctx* ctx, vx_body* bodies, uint32_t n, const float gravity[3], float dt);

/* Backend::findContacts (backend.h:124-126): device broad phase + narrow phase.
 * `out` receives *count contacts in canonical order (pair key, then point);
 * VXHIP_ERR_CAPACITY with *count set to the required size when cap is short. */
int vxhip_find_contacts(vxhip_ctx* ctx, const vx_body* bodies, uint32_t n, const vx_soup* soup, float dt,
                        vx_contact* out, uint32_t cap, uint32_t* count);

/* Backend::solveVelocities (backend.h:131-134; backend_cuda.cu:569-690):
 * colored sweeps over `contacts`.  With warmStart != 0 the contacts are sorted
 * (std::sort on pair key then point), colored, permuted in place into color
 * order with the impulses they came in with, uploaded and warm started; the
 * solved impulses stay on the device for vxhip_fetch_impulses.  A call with
 * warmStart == 0 reuses that device-resident contact set (m must match) and
 * replays the sweeps with the step and options of the warm-started call, as
 * the reference replays the graph it captured then.  Bodies are uploaded and
 * downloaded on every call. */
int vxhip_solve_velocities(vxhip_ctx* ctx, vx_body* bodies, uint32_t n, vx_contact* contacts, uint32_t m,
                           float dt, int warmStart, const vx_solver_options* options);

/* Backend::advanceSubsteps (backend.h:138-142): all substeps of one step on the
 * device (warm start, 2*I colored sweeps per substep, integration between
 * substeps, joints colour by colour, transforms).  `joints` (nj records, 0
 * for none) are solved on the device when there are at least 64; their
 * impulses, reactions and broken flags are written back in place, and
 * removing broken joints stays the caller's (World::finishBrokenJoints).
 * *advanced = 0 when the request is declined: 1..63 joints or a joint type
 * past VX_JOINT_GEAR (backend_cuda.cu:696-701), or the Adaptive iteration
 * policy (backend_cuda.cu:714-715), which is checked after joints between an
 * asleep and an awake body have woken both (backend_cuda.cu:703-707). */
int vxhip_advance_substeps(vxhip_ctx* ctx, vx_body* bodies, uint32_t n, vx_contact* contacts, uint32_t m,
                           vx_joint* joints, uint32_t nj, const float gravity[3], float h, int substeps,
                           const vx_solver_options* options, int* advanced);

/* Backend::fetchImpulses (backend.h:146): accumulated impulses of the last
 * solve, in the order the contacts were left. */
int vxhip_fetch_impulses(vxhip_ctx* ctx, vx_contact* contacts, uint32_t m);

/* Backend::invalidateCaches (backend.h:148): drop device mirrors. */
int vxhip_invalidate_caches(vxhip_ctx* ctx);

/* Backend::lastVelocityIterations (backend.h:161). */
uint32_t vxhip_last_velocity_iterations(const vxhip_ctx* ctx);

/* ------------------------------------------ device-resident world step -- */

/* Uploads a scene and makes it device-resident: bodies, flat geometry and
 * step configuration.  Replaces the per-stage uploads of the reference's
 * device backend (uploadBodies / uploadMeshes, backend_cuda.cu:790-798, 893-915). */
int vxhip_world_load(vxhip_ctx* ctx, const vx_body* bodies, uint32_t n, const vx_soup* soup,
                     const vxhip_config* config);

/* Batched worlds (no reference counterpart: the batched form of the device
 * step).  After vxhip_world_load, declares that body i belongs to world
 * worldOf[i] < worldCount; bodies of different worlds never interact, and
 * each world evolves exactly as it would loaded alone (the broad phase keys
 * cells by world and pairs oversize bodies only within their world; every
 * later stage is per pair, per body chain or per island).  worldCount = 1
 * returns to a single world.  Reset by vxhip_world_load. */
int vxhip_world_set_worlds(vxhip_ctx* ctx, const uint32_t* worldOf, uint32_t worldCount);

/* Solver tuning (no reference counterpart): islands of more than `limit`
 * contacts are solved colour by colour over the whole device instead of by
 * one workgroup per island batch (default 4096; results are identical either
 * way, only the schedule changes). */
int vxhip_world_set_island_limit(vxhip_ctx* ctx, uint32_t limit);

/* Scheduling knobs of the device step (no reference counterpart).  None of
 * them changes a result: every setting steps a world bit for bit the same,
 * they only choose launch shapes and schedules.  velox_amd/autotune.py
 * measures them per workload on the GPU and writes velox_amd/tuning.json,
 * which the Python host applies when it loads a scene.  Zero fields take the
 * defaults given here. */
typedef struct vxhip_tuning {
    uint32_t islandLimit;        /* islands of more contacts go colour by colour over the device (4096) */
    uint32_t batchThreads;       /* 256 or 512: threads of an island-batch workgroup, rows per pass = half (256) */
    uint32_t persistentBatches;  /* 1: island-batch sweeps as one workgroup per CU taking batches from a queue;
                                    2: one workgroup per batch (1) */
    uint32_t longestFirst;       /* 1: batches handed out by pass count, longest first; 2: batch order (1) */
    uint32_t colorLaunchRows;    /* big-island colours with more rows get a launch each, narrower ones run in
                                    single-workgroup runs (1024) */
    uint32_t sideStream;         /* 1: island batches beside the big-island colour chain on a second stream;
                                    2: one stream (1) */
    uint32_t graphSubsteps;      /* 1: substeps 2..S captured once per step in a hipGraph and replayed (when the
                                    step has no big island, joints or recording); 2: launched one by one (2) */
    uint32_t jointPackMax;       /* joint packs (whole components of the joint graph, solved by one workgroup
                                    each) above this many joints take a launch per colour instead (4096) */
} vxhip_tuning;

int vxhip_world_set_tuning(vxhip_ctx* ctx, const vxhip_tuning* tuning);
/* The knobs in effect (defaults filled in). */
int vxhip_world_get_tuning(vxhip_ctx* ctx, vxhip_tuning* out);

/* Advances the resident world `steps` times by dt: the device form of
 * World::stepImpl's physics (world.cpp:3975-4521 without joints, soft bodies
 * or host callbacks): orientation renormalisation, integration, broad phase,
 * narrow phase, wake pass, warm-start matching, greedy coloring, substepped
 * colored TGS solve, conservative-advancement net, positional correction,
 * contact events, sleeping (when enabled), force clearing.  Asynchronous with
 * respect to the host. */
int vxhip_world_step(vxhip_ctx* ctx, float dt, uint32_t steps);

/* Waits for queued device work. */
int vxhip_world_sync(vxhip_ctx* ctx);

/* Copies resident body records to the host (World::body / bodyState). */
int vxhip_world_download_bodies(vxhip_ctx* ctx, vx_body* out, uint32_t n);

/* Replaces resident body records (snapshot restore: invalidateCaches + upload). */
int vxhip_world_upload_bodies(vxhip_ctx* ctx, const vx_body* in, uint32_t n);

/* Contacts solved by the last step, in color order (World's contacts_). */
int vxhip_world_download_contacts(vxhip_ctx* ctx, vx_contact* out, uint32_t cap, uint32_t* count);

/* Contact events of the last step (World::contactEvents, world.cpp:4399-4485):
 * one Begin or Persist event per active pair in pair-key order, then End
 * events for pairs active in the previous step only, in that step's key order.
 * Bodies are dense indices; the host maps them to BodyId handles.
 * VXHIP_ERR_CAPACITY with *count set when cap is short. */
int vxhip_world_download_events(vxhip_ctx* ctx, vx_contact_event* out, uint32_t cap, uint32_t* count);

/* Enables (default) or disables contact events in vxhip_world_step; either
 * call restarts the Begin/Persist history. */
int vxhip_world_set_events(vxhip_ctx* ctx, int enable);

/* Counters of the last step (StepStats subset, world.h:175-208). */
int vxhip_world_stats(vxhip_ctx* ctx, vxhip_stats* out);

/* Joints of the resident world (velox::Joint records, dense body indices).
 * Like the reference's CUDA backend (backend_cuda.cu:691-761), the device
 * solves them only when there are at least 64: fewer fail with
 * VXHIP_ERR_UNAVAILABLE and stay with the host's World::solveJoints.  Each
 * vxhip_world_step then drops contacts of joint-connected pairs
 * (world.cpp:4045-4066), wakes sleeping bodies jointed to awake ones, solves
 * the joints in colour order after the contacts of every substep
 * (joint_solver.h, 8 passes), and removes broken joints as
 * World::finishBrokenJoints does (world.cpp:3765-3783): from the end, each
 * swapped with the last.  vxhip_world_load clears the joints. */
int vxhip_world_set_joints(vxhip_ctx* ctx, const vx_joint* joints, uint32_t count);
/* The same for any number of joints (no reference counterpart): every joint
 * set, fewer than 64 included, is solved by the colour-ordered device path
 * above, where the reference's World would hand a set of fewer than 64 to its
 * host solver.  Batched worlds take their joints this way: the batch's joint
 * set is large while each world's own is small, and a world stepped alone
 * must step exactly as it does inside the batch. */
int vxhip_world_set_joints_all(vxhip_ctx* ctx, const vx_joint* joints, uint32_t count);

/* The joints after the last step (after removals). */
int vxhip_world_download_joints(vxhip_ctx* ctx, vx_joint* out, uint32_t cap, uint32_t* count);

/* Joints removed by the last step, in removal order (World::jointBreakEvents,
 * world.cpp:3776-3780): the joint's dense index when removed, its bodies,
 * reaction force and torque. */
typedef struct vx_joint_break {
    uint32_t index, a, b;
    float force, torque;
} vx_joint_break;
int vxhip_world_download_joint_breaks(vxhip_ctx* ctx, vx_joint_break* out, uint32_t cap, uint32_t* count);

/* Multi-TOI scheduler settings (WorldMultiToiSettings, ccd.h:111-122; defaults
 * 1, 4, 0.1, 1e-3, 256, 1).  MotionQuality::High bodies that touch only static
 * or other such bodies are replayed by processHighToi (world.cpp:3019-3142)
 * inside vxhip_world_step.  Invalid settings (validateMultiToiSettings,
 * world.cpp:60-67) fail with VXHIP_ERR_INVALID.  Replaces
 * World::setMultiToiSettings (world.cpp:580-584) for the device step. */
typedef struct vx_multi_toi_settings {
    uint8_t enabled;                 /* CcdConfig::enabled */
    uint8_t enableSubstepSplitting;
    uint8_t reserved[2];
    uint32_t maxToiEventsPerBody;
    float toiVelocityFloor;
    float toiPenetrationBias;
    uint32_t maxTotalEventsPerStep;
} vx_multi_toi_settings;
int vxhip_world_set_multi_toi(vxhip_ctx* ctx, const vx_multi_toi_settings* settings);

/* One time-of-impact hit (MultiToiHit, world.h:121-127); `body` is the dense
 * index of the body hit. */
typedef struct vx_toi_hit {
    float toi;
    uint32_t body;
    float point[3];
    float normal[3];   /* from the hit body toward the queried body */
    float fraction;    /* toi / dt */
} vx_toi_hit;

/* World::queryMultiToi (world.h:405, world.cpp:3146-3185) on the resident
 * world: findToi of body `body` against every other body that passes the
 * collision filter, with the multi-TOI settings' velocity floor and
 * penetration bias.  *count receives the number of hits; `out` receives the
 * first min(cap, *count) of them sorted by (toi, body index).  The caller
 * applies the reference's cap (min of maxToiEventsPerBody and
 * maxTotalEventsPerStep) and, for a World whose handles are not in index
 * order, re-sorts ties by handle before capping.  A World with a collision
 * filter callback keeps the host query (the callback is host code).  A
 * non-finite or non-positive dt fails with VXHIP_ERR_INVALID, as the
 * reference throws; batched contexts fail with VXHIP_ERR_UNAVAILABLE. */
int vxhip_world_query_multi_toi(vxhip_ctx* ctx, uint32_t body, float dt, vx_toi_hit* out, uint32_t cap,
                                uint32_t* count);

/* Moves the world origin: every stored world coordinate of the resident world
 * (body positions, plane offsets, mesh vertices, mesh and BVH bounds, the last
 * step's contacts, the warm-start cache, the last events' points) becomes
 * value - offset.  Replaces World::shiftOrigin (world.cpp:1978-2036): a
 * non-finite offset or result fails with VXHIP_ERR_INVALID and changes
 * nothing. */
int vxhip_world_shift_origin(vxhip_ctx* ctx, const float offset[3]);

/* The device state the next vxhip_world_step reads, as one host blob: bodies,
 * warm-start cache, the events' pair history, the last events and contacts.
 * The device form of World::saveSnapshot / restoreSnapshot
 * (world.cpp:2038-2127); as there, sleep-manager counters are not part of
 * it.  Restoring needs a world of the same body count (VXHIP_ERR_INVALID
 * otherwise); a step after a restore continues bit for bit as after the save. */
int vxhip_world_state_size(vxhip_ctx* ctx, size_t* bytes);
int vxhip_world_save_state(vxhip_ctx* ctx, void* out, size_t cap);
int vxhip_world_restore_state(vxhip_ctx* ctx, const void* in, size_t size);

/* Island sleeping (SleepConfig, sleep.h:98-160; field meaning and defaults as
 * there: 0.05, 0.05, 0.01, 0.5, 0.25, 0.25, 1, 1, 0.02, 10). */
typedef struct vx_sleep_config {
    float linearVelocityThreshold;
    float angularVelocityThreshold;
    float accelerationThreshold;
    float timeToSleep;
    float timeToDrowsy;
    float drowsySimulationRate;  /* validated, not applied (as in the reference) */
    float contactStabilityThreshold;
    int32_t contactStabilityFrames;
    uint8_t enableGradualSleep;
    uint8_t enableContactStability;
    uint8_t reserved[2];
} vx_sleep_config;

/* SleepStats (sleep.h:167-180) without the wall-clock field. */
typedef struct vx_sleep_stats {
    uint32_t totalDynamicBodies, awakeBodies, drowsyBodies, sleepingBodies;
    uint32_t islandCount, sleepingIslandCount, drowsyIslandCount, awakeIslandCount;
    uint64_t totalSleepTransitions, totalWakeTransitions, totalDrowsyTransitions;
} vx_sleep_stats;

/* Turns the device sleep update on with `config` (validated and clamped as
 * SleepConfig::validate, sleep.h:143-159), or off with NULL (the default).
 * While on, every vxhip_world_step ends like World::stepImpl
 * (world.cpp:4491-4521): SleepManager::update (sleep.cpp:50-325) over the
 * step's contacts, then the warm-start cache with the contacts of inert pairs
 * carried over; the wake pass (world.cpp:4109-4128) runs whenever a body is
 * asleep.  Contact pair history for events is kept either way.  Resets the
 * per-body contact stability counters and the cumulative transition counts.
 * Replaces World::setSleepConfig (world.h:741) for the device step.
 * Limited to worlds of at most 256 bodies in this build (VXHIP_ERR_CAPACITY
 * above, here and in vxhip_world_load while sleeping is on): a 1,001-body
 * world with sleeping on faulted the device and the cause is not yet found. */
int vxhip_world_set_sleep(vxhip_ctx* ctx, const vx_sleep_config* config);

/* Sleep statistics after the last step (World::sleepStats, world.h:744). */
int vxhip_world_sleep_stats(vxhip_ctx* ctx, vx_sleep_stats* out);

/* Stage timings measured with HIP events on the context's own stream while
 * profiling is enabled (accumulated over steps; StepStats timings,
 * world.cpp:4523-4575).  Every launch of the kernels below is bracketed by
 * its own event pair, giving each one's device time, launch count and the
 * work units it processed, indexed by VXHIP_KERNEL_*: contact visits
 * (contacts x sweeps) for the solve kernels, candidate pairs for the narrow
 * phase and the CCD check, walked pairs for the CCD dataflow, bodies for the
 * pair emission. */
enum {
    VXHIP_KERNEL_BATCH = 0,        /* kBatchSweeps, velocity sweeps (island batches in LDS) */
    VXHIP_KERNEL_BATCH_WARM = 1,   /* kBatchSweeps, warm start */
    VXHIP_KERNEL_ISLAND = 2,       /* kIslandSweeps (one larger island per workgroup) */
    VXHIP_KERNEL_COLOR = 3,        /* big-island colours: kSweepColorPair, kSweepColorRun, kSweepTailPair */
    VXHIP_KERNEL_NARROW = 4,       /* kNarrow (units: candidate pairs) */
    VXHIP_KERNEL_CCD_CHECK = 5,    /* kCcdCheck (units: pairs checked) */
    VXHIP_KERNEL_CCD_DATAFLOW = 6, /* kCcdDataflow (units: pairs walked in order) */
    VXHIP_KERNEL_GRID_PAIRS = 7,   /* kGridPairs (units: bodies) */
    VXHIP_KERNEL_COUNT = 8
};
typedef struct {
    uint32_t steps;
    float beginMs, broadMs, narrowMs, matchMs, colorMs, solveMs, ccdMs, positionMs, finishMs;
    float kernelMs[VXHIP_KERNEL_COUNT];
    uint32_t kernelLaunches[VXHIP_KERNEL_COUNT];
    uint64_t kernelVisits[VXHIP_KERNEL_COUNT];
} vxhip_timings;

/* Enables (non-zero) or disables per-stage event timing and resets it. */
int vxhip_world_profile(vxhip_ctx* ctx, int enable);
int vxhip_world_timings(vxhip_ctx* ctx, vxhip_timings* out);

/* ------------------------------------------------- differentiable form -- */
/* Recording of the frozen active set: while on, each vxhip_world_step keeps
 * its contacts in solve order with the warm-start impulses they started
 * from, and for each the previous step's solve position whose final
 * impulses it took (0xffffffff: none); vxhip_world_download_record copies
 * the last step's (count first; nothing is copied when cap is too small).
 * No reference counterpart: the record is the differentiable form's. */
int vxhip_world_record(vxhip_ctx* ctx, int enable);
int vxhip_world_download_record(vxhip_ctx* ctx, vx_contact* contacts, uint32_t* match, uint32_t cap, uint32_t* count);


/* Gradient rollouts of the resident world (no reference counterpart: the
 * differentiable form is this build's own).  The rollout is the ordinary
 * device step; its gradients come from hand-written adjoint kernels of the
 * substepped contact solve, the transform advance and the positional
 * correction, with the active set frozen per step (each step's contacts,
 * their geometry, colours and warm-start matches are the rollout's; the
 * branches the solve took are replayed).  A checkpoint of the state a step
 * reads from the one before is kept on the device every `every` steps; the
 * backward replays one segment at a time.  Admitted worlds: dynamic bodies of
 * isotropic inertia (spheres, cubes), geometric-mean friction without rolling
 * or spinning resistance, no sleeping, joints or multi-TOI bodies, the
 * two-axis friction model, each batched world's bodies contiguous
 * (VXHIP_ERR_UNAVAILABLE otherwise).
 *
 * vxhip_diff_forward: `steps` device steps from the current state; `loss`
 * (one per world, may be NULL) receives |c_w - t_w|^2 for the mass-weighted
 * centre c_w of world w's dynamic bodies after the last step and the target
 * t_w = targets[3w..3w+2]. */
int vxhip_diff_forward(vxhip_ctx* ctx, float dt, uint32_t steps, uint32_t every, const float* targets, float* loss);

typedef struct {
    uint32_t steps, checkpoints, worlds;
    uint32_t ccdRescued;      /* pairs the CCD safety net rescued during the rollout: not differentiated */
    uint32_t replayMismatch;  /* body-steps whose tape replay differed from the rollout (0: bit for bit) */
    uint32_t maxContacts;     /* largest contact count of a step */
} vxhip_diff_stats;

/* Gradients of sum_w lossWeight[w] * loss_w with respect to the rollout's
 * initial linear and angular velocities (gV, gW: 3 per body), friction
 * coefficients (gFriction) and masses (gMass: through the inverse mass, the
 * inverse inertia it scales and the centres of mass), for every body (zero
 * for static ones).  Leaves the world at the rollout's end state with an
 * empty warm-start cache.  Any output may be NULL. */
int vxhip_diff_backward(vxhip_ctx* ctx, const float* lossWeight, float* gV, float* gW, float* gFriction, float* gMass,
                        vxhip_diff_stats* stats);

/* ------------------------------------------------------- stage probes -- */
/* Single stages on caller data, for parity checks of one stage at a time. */

/* Broad phase only: candidate pairs as sorted contact keys (a<<32|b, convex
 * body first for plane/mesh pairs). */
int vxhip_probe_pairs(vxhip_ctx* ctx, const vx_body* bodies, uint32_t n, const vx_soup* soup, float dt,
                      uint64_t* out, uint32_t cap, uint32_t* count);

/* Greedy coloring of contacts in the given order (backend_cuda.cu:597-630). */
int vxhip_probe_color(vxhip_ctx* ctx, const vx_body* bodies, uint32_t n, const vx_contact* contacts, uint32_t m,
                      uint32_t* colors, uint32_t* numColors);

/* Warm-start matching of `cur` against `prev` (sorted by pair key), in place
 * (world.cpp:4133-4184). */
int vxhip_probe_warm_start(vxhip_ctx* ctx, vx_contact* cur, uint32_t m, const vx_contact* prev, uint32_t mp);

/* The device step's stable LSD radix sort of (key, value) pairs by the low
 * `bits` (0..64) of the key, in place (no reference counterpart: it stands
 * where the reference calls Thrust's sort, and any stable sort gives its
 * result, so it is checked against a host stable sort). */
int vxhip_probe_radix_sort(vxhip_ctx* ctx, uint64_t* keys, uint32_t* vals, uint32_t n, int bits);

/* std::sort (libstdc++) of (key, value) pairs by key, reproduced exactly on
 * the device, in place: the order equal keys are left in is std::sort's.
 * With segStart (segments + 1 non-decreasing offsets, 0 to n), each segment
 * is sorted as its own std::sort call would sort it.  The device step orders
 * its warm-start cache this way, as World::step does with std::sort
 * (world.cpp:4518-4521). */
int vxhip_probe_std_sort(vxhip_ctx* ctx, uint64_t* keys, uint32_t* vals, uint32_t n, const uint32_t* segStart,
                         uint32_t segments);

/* The same sort on caller-owned device memory, allocating nothing (no
 * reference counterpart: for callers that cannot allocate, such as code
 * under HIP graph capture or tight serving loops).  keys, vals and segStart
 * are device pointers; the working memory is a fixed 4 MiB device array per
 * GPU, so n <= 262,144 and segments <= 4,096 (VXHIP_ERR_CAPACITY beyond).
 * Segment offsets that do not rise from 0 to n fail with VXHIP_ERR_INVALID
 * and leave the keys unsorted.  Runs on the context's stream and returns
 * when the sort is done.  Calls that overlap on one GPU share the array and
 * race; the context must belong to the current device. */
int vxhip_std_sort_noalloc(vxhip_ctx* ctx, uint64_t* keys, uint32_t* vals, uint32_t n, const uint32_t* segStart,
                           uint32_t segments);

/* ------------------------------------------------------ synthetic scenes -- */
/* Host utility: builds body records and flat geometry for the synthetic
 * workloads the device step is measured on (no reference counterpart; the
 * reference builds scenes through World::addBox/addSphere/addCapsule/
 * addConvexHull/addStaticPlane/addStaticHeightfield, world.cpp:741-1477).
 * Names: "lattice" (param = boxes per side), "debris" (param = body count),
 * "pile", "terrain", "freeflight". */
typedef struct vxhip_scene vxhip_scene;
int vxhip_scene_build(const char* name, uint32_t param, vxhip_scene** out);
uint32_t vxhip_scene_body_count(const vxhip_scene* scene);
const vx_body* vxhip_scene_bodies(const vxhip_scene* scene);
const uint32_t* vxhip_scene_worlds(const vxhip_scene* scene, uint32_t* worldCount);
int vxhip_scene_soup(const vxhip_scene* scene, vx_soup* out);
/* The scene's joints (ball joints of the batched scene's capsule chains), or
 * null with *count = 0. */
const vx_joint* vxhip_scene_joints(const vxhip_scene* scene, uint32_t* count);
int vxhip_scene_config(const vxhip_scene* scene, vxhip_config* config, float* dt);
void vxhip_scene_free(vxhip_scene* scene);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* VXHIP_H */
