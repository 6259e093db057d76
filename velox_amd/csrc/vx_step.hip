// Host orchestration and C ABI of libvxhip: the device step of Velox on one
// MI355X (gfx950), device-resident across steps.  See include/vxhip.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <string>
#include <vector>
#include "vx_ccd.hpp"
#include "vx_solve.hpp"
#include "vx_sleep.hpp"
#include "vx_toi.hpp"
#include "vx_joints.hpp"
#include "vx_diff.hpp"
#include "vx_internal.hpp"
#include "../../include/vxhip.h"

using namespace vxk;

namespace {

thread_local std::string gLastError = "no error";

struct Failure {
    int code;
};

int classify(hipError_t e) {  // backend_cuda.cu:24-40
    switch (e) {
    case hipErrorOutOfMemory: return VXHIP_ERR_MEMORY;
    case hipErrorNoDevice:
    case hipErrorLaunchFailure:
    case hipErrorLaunchTimeOut:
    case hipErrorUnknown:
    case hipErrorNotInitialized:
    case hipErrorIllegalAddress: return VXHIP_ERR_DEVICE_LOST;
    default: return VXHIP_ERR_INVALID;
    }
}

[[noreturn]] void fail(int code, const std::string& what) {
    gLastError = what;
    throw Failure{code};
}

void check(hipError_t e, const char* expr, int line) {
    if (e == hipSuccess) return;
    char buf[512];
    std::snprintf(buf, sizeof(buf), "vxhip: %s (%s) at vx_step.hip:%d", hipGetErrorString(e), expr, line);
    fail(classify(e), buf);
}
#define HIPCK(x) check((x), #x, __LINE__)

// Host-forced fault hook (backend_cuda.cu:69-78).
void injectFailureForTesting() {
    const char* mode = std::getenv("VELOX_HIP_TEST_FAIL");
    if (!mode || !mode[0]) return;
    if (!std::strcmp(mode, "oom")) fail(VXHIP_ERR_MEMORY, "vxhip test injection: allocation failure");
    if (!std::strcmp(mode, "device-loss")) fail(VXHIP_ERR_DEVICE_LOST, "vxhip test injection: device loss");
}

template <class T>
struct Buf {
    T* p = nullptr;
    size_t cap = 0;
    void reserve(size_t n) {
        if (n <= cap) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = n + n / 4 + 64;
        HIPCK(hipMalloc((void**)&p, want * sizeof(T)));
        cap = want;
    }
    // Like reserve, but the first `keep` elements survive a reallocation.
    void grow(size_t n, size_t keep, hipStream_t s) {
        if (n <= cap) return;
        T* old = p;
        p = nullptr;
        cap = 0;
        size_t want = n + n / 4 + 64;
        HIPCK(hipMalloc((void**)&p, want * sizeof(T)));
        cap = want;
        if (old && keep) HIPCK(hipMemcpyAsync(p, old, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
        if (old) {
            HIPCK(hipStreamSynchronize(s));
            (void)hipFree(old);
        }
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { release(); }
};

inline uint32_t blocks(uint64_t n, int t = kThreads) { return (uint32_t)((n + t - 1) / t); }

// One step of a differentiable rollout as its adjoint replays it.
struct DiffStepRec {
    uint32_t m = 0, prevCount = 0, ccdRescued = 0;
    Buf<vx_body> start;     // bodies at the step's start
    Buf<vx_contact> cs;     // contacts in solve order, with their warm-start impulses
    Buf<uint32_t> match;    // per contact: the previous step's solve position it took impulses from, or ~0
    Buf<uint32_t> colorOf;  // per contact: colour
    Buf<uint32_t> matchCanon;
};

// The state a step reads from the one before it: bodies and warm-start cache.
struct DiffCheckpoint {
    uint32_t prevCount = 0;
    bool prevExact = false;
    Buf<vx_body> bodies;
    Buf<vx_contact> prev;
    Buf<uint32_t> prevRank;
};

}  // namespace

struct vxhip_ctx {
    hipStream_t stream = nullptr;
    int device = 0;  // the device current at vxhip_create
    // resident scene
    uint32_t n = 0;
    Buf<vx_body> bodies;
    Buf<vx_vec3> vertices, hullPts;
    Buf<uint32_t> indices, triRefs, hullFaces;
    Buf<vx_mesh> meshes;
    Buf<vx_bvh_node> nodes;
    Buf<vx_compound_child> children;
    Soup soup;
    uint32_t vertexCount = 0, meshCount = 0, nodeCount = 0;
    size_t soupSig[8] = {};
    vxhip_config cfg{};
    bool anyAsleep = false;
    // per-step scratch
    Buf<vx_vec3> prevPos;
    Buf<vx_quat> prevQ;
    Buf<float4> lo, hi;
    Buf<uint64_t> keys, keysTmp;
    Buf<uint32_t> idx, idxTmp, oversize;
    Buf<uint64_t> pairs;
    Buf<uint64_t> pairClass, pairClassTmp;  // candidate pairs by shape class (kPairClass)
    Buf<uint32_t> pairOrder, pairOrderTmp;
    Buf<uint32_t> pairVals, counts, offsets;
    Buf<vx_contact> slots, contacts, solve, prev;
    Buf<uint32_t> colorVals, colorVals2, pred, color, colorCount, order, colorStart;
    Buf<unsigned long long> maskAfter;
    Buf<int> nextAfter;
    Buf<uint8_t> ccdFlag;
    Buf<uint32_t> ccdRel, ccdRelPos, ccdList, ccdFirstRel;
    // Batched worlds: world id per body (bodies of different worlds never
    // interact) and the per-world oversize lists of the broad phase.
    Buf<uint32_t> worldOf, worldOversize, worldStart, worldFill, oversizeByWorld;
    uint32_t worldCount = 1, worldN = 0;
    bool worldsContiguous = false;  // bodies lie world after world (non-decreasing world ids)
    Buf<uint32_t> colorSeg, colorSegCount;
    Buf<uint2> cellRange;  // per-cell run of the sorted keys (small grids)
    // Contact events of the last step and the pair keys they leave for the next.
    Buf<uint64_t> evKeys, evKeysTmp, evPrevKeys, evCurKeys;
    Buf<uint32_t> evVals, evValsTmp, evHead, evRunPos, evEndFlag, evEndPos;
    Buf<vx_contact_event> events;
    uint32_t eventCount = 0, evPrevCount = 0;
    bool eventsEnabled = true;
    Buf<uint64_t> evAppKeys, evMerged;
    Buf<uint32_t> evEndP, evEndBits, evApp, evEmit;
    // Sleeping (off unless configured) and the warm-start cache it needs.
    bool sleepOn = false;
    vxk::SleepParams sleepParams{};
    vx_sleep_config sleepCfg{};
    Buf<uint32_t> slParent, slFlags, slIslandMin, slRootState;
    Buf<uint16_t> slStability;
    vx_sleep_stats sleepStats{};
    Buf<vx_contact> prevStage, prevNext;
    Buf<uint32_t> carry, carryRank, cacheVals, cacheValsTmp;
    Buf<uint64_t> cacheKeys, cacheKeysTmp;
    // The warm-start cache is matched in the order std::sort leaves it.
    // prevExact: `prev` is in that order; otherwise it is canonical and
    // prevRank holds each entry's position in the solve order it came from.
    bool prevExact = false;
    Buf<uint32_t> prevRank;
    Buf<uint64_t> exWorld, exWorldTmp, exKeys;
    Buf<uint32_t> exIdx, exIdxTmp, exVals, exCount, exStart;
    Buf<float4> sVel, sAng, sInvI, sFrame, sRows, sImp;
    Buf<uint4> sMeta, sMetaK;
    Buf<float4> sSide;
    Buf<uint32_t> solvePos;
    // island partition of the colour-ordered contacts (one step)
    Buf<uint32_t> ldsSize, ldsOffset, globalFlag, globalRank, bodyRoot, bodyRank, bodyCount, bodyOffset, bodyList,
        slotOf, rangeStart, rangeEnd, slotPair;
    Buf<uint2> batchBodies;
    Buf<uint32_t> uf, islandOf, islandSize, bigFlag, bigPos, bigIdx, bigColorCount, bigColorStart, colorSolve;
    Buf<uint64_t> smallKeys, smallKeysTmp;
    Buf<uint32_t> smallIdx, smallIdxTmp, runHead, islandHead, runPos, islandPos, runStart, islandFirstRun;
    std::vector<uint32_t> bigColorStartHost;
    uint32_t smallCount = 0, bigCount = 0, islandCount = 0, batchCount = 0, ldsBatchCount = 0, ldsContactCount = 0,
             largestIsland = 0,
             maxBatchPasses = 0;
    uint32_t islandLimit = 4096;
    // Scheduling knobs (vxhip_tuning); none changes a result.
    uint32_t batchThreads = 256, colorLaunchRows = 1024;
    bool persistentBatches = true, longestFirst = true, sideOn = true, graphSubsteps = false;
    uint32_t cuCount = 256;
    // Island-batch queue (slot counter, finished workgroups) and the batches'
    // longest-first order of the current step.
    Buf<uint32_t> batchQueue, batchOrder, batchOrderTmp;
    Buf<uint64_t> batchKeys, batchKeysTmp;
    bool batchOrderReady = false;
    bool batchSmall = false;  // this step's LDS batches all fit kBatchSweeps' small layout
    // Substeps 2..S of a step replayed from one captured graph.
    hipGraphExec_t subExec = nullptr;
    uint32_t graphReplays = 0, graphCaptures = 0;
    uint32_t graphLaunches[VXHIP_LAUNCH_KINDS] = {};  // launches of each kind one replay stands for
    // LDS tail run of the big-island colours (colours tailColor..K-1).
    uint32_t tailColor = 0;
    Buf<uint64_t> tailKeys, tailKeysTmp;
    Buf<uint32_t> tailVals, tailValsTmp, tailHead, tailPos, tailRowSlot, tailSlotBody;
    // Joints (solved on the device when there are at least 64, like the
    // reference's CUDA backend): records in the caller's order, the colour
    // order of their solve, and the pairs they keep from colliding.
    Buf<vx_joint> joints;
    uint32_t jointCount = 0, jointColors = 0, jointExcludedCount = 0, jointBreakCount = 0;
    std::vector<vx_joint> hostJoints;
    std::vector<uint8_t> hostDynamic;
    std::vector<uint32_t> jointColorStartHost;
    Buf<uint32_t> jointOrder, jointKeep, jointKeepPos;
    // Joint packs: joints whose dynamic bodies connect them (the components of
    // the joint graph), concatenated into packs of about kJointPack joints,
    // each solved colour by colour in one workgroup.  Packs above
    // jointPackMax joints (a tuning knob) go through the per-colour launches
    // (jointOrder).
    uint32_t jointPacks = 0, jointPackMax = kJointPackMax;
    Buf<uint32_t> packOrder, packRunStart, packFirstRun;
    Buf<uint64_t> jointExcluded;
    Buf<vx_contact> contactsTmp;
    Buf<vxk::JointBreak> jointBreaks;
    // Multi-TOI scheduler for MotionQuality::High bodies.
    bool anyHigh = false, toiProcessed = false;
    vxk::ToiSettings toiSettings{1u, 1u, 4u, 256u, 0.1f, 1e-3f};  // ccd.h:111-122 defaults
    Buf<vx_body> toiStarts;
    Buf<uint8_t> toiEligible;
    Buf<uint32_t> toiE, toiS, toiEvents, toiBits;
    Buf<float4> toiHit;
    Buf<vxk::ToiState> toiState;
    uint32_t toiEventCount = 0;
    Buf<uint32_t> ccdDone;
    Buf<uint32_t> toiQueryBits;  // vxhip_world_query_multi_toi: per other body
    Buf<float4> toiQueryHit;
    Buf<float4> ccdGap;
    Buf<uint8_t> temp;
    Buf<uint32_t> counters;  // device counters, see Counter
    Buf<uint32_t> extHist;   // body extents by bucket (kAabb), cleared by kCellSize
    uint32_t* hostCounters = nullptr;
    uint32_t pairCap = 0;
    uint32_t pairsCount = 0;
    uint32_t contactCount = 0;
    uint32_t prevCount = 0;
    uint32_t solveCount = 0;
    float solveH = 0.0f;               // step of the last warm-started vxhip_solve_velocities
    vx_solver_options solveOptions{};  // and its options, reused by the calls after it
    uint32_t numColors = 0;
    std::vector<uint32_t> colorStartHost;
    vxhip_stats stats{};
    uint32_t lastIterations = 0;
    bool loaded = false;
    // profiling (vxhip_world_profile): events on this context's stream
    bool profile = false;
    // Island batches run beside the big-island colour chain (disjoint bodies
    // and contacts) on the device's shared side stream, forked and joined
    // with these events.
    hipEvent_t forkEv = nullptr, joinEv = nullptr;
    std::vector<hipEvent_t> stageEv;     // one per stage boundary
    std::vector<hipEvent_t> launchEv;    // pairs around colour-solve launches
    uint32_t launchUsed = 0;
    std::vector<uint64_t> launchContacts;
    std::vector<int> launchKind;
    vxhip_timings timings{};
    // Differentiable rollouts (vxhip_diff_*): the step being recorded, the
    // checkpoints of the last rollout and the adjoint's buffers.
    DiffStepRec* diffCap = nullptr;
    bool recordOn = false;  // vxhip_world_record: the last step's frozen data kept for download
    DiffStepRec record;
    uint32_t diffSteps = 0, diffEvery = 8, diffW = 0;
    float diffDt = 0.0f;
    bool diffReady = false;
    std::vector<std::unique_ptr<DiffCheckpoint>> diffCps;
    std::vector<std::unique_ptr<DiffStepRec>> diffSeg;
    std::vector<uint32_t> diffBodyStartHost;
    Buf<uint32_t> diffBodyStart;
    Buf<float> diffTarget, diffLoss, diffLossBar, diffAdjMass, diffG;
    Buf<vx_body> diffFinal, diffWork, diffPost;
    Buf<vx_body> diffSub[8], diffPre[8];  // per substep: bodies before the solve, and before the transform
    Buf<float4> diffRows[8], diffTape[8], diffImp0;
    Buf<float2> diffCTape;
    Buf<float4> diffAdjX, diffAdjQ, diffAdjV, diffAdjW, diffAdjParam, diffLamBar, diffLamPrev, diffRowBar;
    Buf<float4> dVel, dAng, dInvI, dFrame, dImp, dSide;
    Buf<uint4> dMeta;
    Buf<uint32_t> dIota, dList, dListTmp, dWStart, dWCount;
    Buf<uint64_t> dWKeys, dWKeysTmp;
    Buf<uint32_t> diffOverflow;
};

namespace {

enum Counter : uint32_t {
    cExtent = 0, cBounds = 1 /* 6 ints */, cOversize = 7, cPairs = 8, cContacts = 9, cEntries = 10,
    cColors = 11, cColorError = 12, cCcdFlagged = 13, cCcdFirst = 14, cCcdExamined = 15, cCcdRescued = 16,
    cOversize2 = 17, cSmall = 18, cBig = 19, cRuns = 20, cIslands = 21, cBatches = 22, cLdsBatches = 23, cRanges = 24, cLargest = 28, cPasses = 29, cBatchError = 30, cCcdRelevant = 33, cLdsContacts = 34, cEvActive = 35, cEvRuns = 36, cEvEnds = 37, cEvInert = 38, cEvCount = 39,
    cSlDynamic = 40 /* dynamic, awake, drowsy, asleep, to sleep, to drowsy */, cSlWoke = 46,
    cSlIslands = 47 /* islands, asleep, drowsy, awake */, cCarried = 51, cTailSlots = 52, cShiftBad = 53, cJointsLeft = 54, cJointBreaks = 55,
    cCcdRound = 56 /* pairs each closure round added, 8 rounds */, cWarmTies = 64, cBatchMaxBodies = 72, cBatchMaxRows = 73, cJointsBroken = 74, cCount = 75
};
static_assert(cCount <= 128, "kResetCounters runs one 128-lane workgroup");

void* tempFor(vxhip_ctx* c, size_t bytes) {
    c->temp.reserve(bytes);
    return c->temp.p;
}

void readCounters(vxhip_ctx* c) {
    HIPCK(hipMemcpyAsync(c->hostCounters, c->counters.p, cCount * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
}

enum Stage : int { sBegin = 0, sBroad, sNarrow, sMatch, sColor, sSolve, sCcd, sPosition, sFinish, sEnd, sCount };

void markStage(vxhip_ctx* c, int stage) {
    if (c->profile) HIPCK(hipEventRecord(c->stageEv[stage], c->stream));
}

// Events bracket every launch of the event-timed kernels while profiling.
void solveLaunchBegin(vxhip_ctx* c) {
    if (!c->profile) return;
    if (c->launchUsed + 2 > c->launchEv.size()) {
        size_t old = c->launchEv.size();
        c->launchEv.resize(old + 4096);
        for (size_t i = old; i < c->launchEv.size(); ++i) HIPCK(hipEventCreate(&c->launchEv[i]));
    }
    HIPCK(hipEventRecord(c->launchEv[c->launchUsed], c->stream));
}

void solveLaunchEnd(vxhip_ctx* c, int kind, uint64_t contacts) {
    if (!c->profile) return;
    HIPCK(hipEventRecord(c->launchEv[c->launchUsed + 1], c->stream));
    c->launchUsed += 2;
    c->launchContacts.push_back(contacts);
    c->launchKind.push_back(kind);
}

void collectProfile(vxhip_ctx* c) {
    if (!c->profile) return;
    HIPCK(hipStreamSynchronize(c->stream));
    float* stageMs[sCount - 1] = {&c->timings.beginMs, &c->timings.broadMs, &c->timings.narrowMs, &c->timings.matchMs,
                                  &c->timings.colorMs, &c->timings.solveMs, &c->timings.ccdMs, &c->timings.positionMs,
                                  &c->timings.finishMs};
    for (int st = 0; st + 1 < sCount; ++st) {
        float ms = 0.0f;
        HIPCK(hipEventElapsedTime(&ms, c->stageEv[st], c->stageEv[st + 1]));
        *stageMs[st] += ms;
    }
    for (uint32_t i = 0; i < c->launchUsed; i += 2) {
        float ms = 0.0f;
        HIPCK(hipEventElapsedTime(&ms, c->launchEv[i], c->launchEv[i + 1]));
        const int kind = c->launchKind[i / 2];
        c->timings.kernelMs[kind] += ms;
        c->timings.kernelLaunches[kind] += 1;
        // The dataflow's relevant pairs are counted on the device (read with
        // the step's last counters).
        c->timings.kernelVisits[kind] +=
            kind == VXHIP_KERNEL_CCD_DATAFLOW ? c->hostCounters[cCcdRelevant] : c->launchContacts[i / 2];
    }
    c->launchUsed = 0;
    c->launchContacts.clear();
    c->launchKind.clear();
    c->timings.steps += 1;
}

// Resets the per-stage counters; the sticky words (cBatchError, cWarmTies)
// are cleared only at the start of a step (keepSticky false).
__global__ void kResetCounters(uint32_t* k, bool keepSticky = true) {
    uint32_t t = threadIdx.x;
    if (t >= cCount || (keepSticky && (t == cBatchError || t == cWarmTies))) return;
    uint32_t v = 0;
    if (t >= cBounds && t < cBounds + 3) v = 0x7fffffffu;       // min cell
    else if (t >= cBounds + 3 && t < cBounds + 6) v = 0x80000000u;  // max cell
    else if (t == cCcdFirst) v = 0xffffffffu;
    k[t] = v;
}

__global__ void kFill32(uint32_t* p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void kIota(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = i;
}
__global__ void kColorOfSolve(const uint32_t* color, const uint32_t* order, uint32_t m, uint32_t* out) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k < m) out[k] = color[order[k]];
}
__global__ void kColorKeys(const uint32_t* color, uint32_t m, uint64_t* keys, uint32_t* vals) {
    uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k < m) { keys[k] = color[k]; vals[k] = k; }
}

void fill32(vxhip_ctx* c, uint32_t* p, uint32_t n, uint32_t v) {
    if (n) kFill32<<<blocks(n), kThreads, 0, c->stream>>>(p, n, v);
}

// Up to kFillMax buffers filled by one launch (each fill is otherwise a
// launch of its own, and the step's fills come in runs).
constexpr int kFillMax = 6;
struct FillSet {
    uint32_t* p[kFillMax];
    uint32_t n[kFillMax];
    uint32_t v[kFillMax];
    int count;
};
__global__ void kFillMany(FillSet f) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
#pragma unroll
    for (int q = 0; q < kFillMax; ++q)
        if (q < f.count && i < f.n[q]) f.p[q][i] = f.v[q];
}
struct Fill { uint32_t* p; uint32_t n; uint32_t v; };
void fillMany(vxhip_ctx* c, std::initializer_list<Fill> fills) {
    FillSet f{};
    uint32_t most = 0;
    for (const Fill& x : fills) {
        if (!x.n) continue;
        if (f.count == kFillMax) fail(VXHIP_ERR_INVALID, "vxhip: too many fills in one launch");
        f.p[f.count] = x.p; f.n[f.count] = x.n; f.v[f.count] = x.v;
        ++f.count;
        most = std::max(most, x.n);
    }
    if (f.count) kFillMany<<<blocks(most), kThreads, 0, c->stream>>>(f);
}

template <class T>
void upload(vxhip_ctx* c, Buf<T>& b, const T* host, uint32_t count) {
    b.reserve(count ? count : 1);
    if (count) HIPCK(hipMemcpyAsync(b.p, host, (size_t)count * sizeof(T), hipMemcpyHostToDevice, c->stream));
}

void uploadSoup(vxhip_ctx* c, const vx_soup* s) {
    vx_soup e{};
    if (!s) s = &e;
    upload(c, c->vertices, s->vertices, s->vertexCount);
    upload(c, c->indices, s->indices, s->indexCount);
    upload(c, c->meshes, s->meshes, s->meshCount);
    upload(c, c->nodes, s->bvhNodes, s->bvhNodeCount);
    upload(c, c->triRefs, s->bvhTriRefs, s->bvhTriRefCount);
    upload(c, c->hullPts, s->hullPoints, s->hullPointCount);
    upload(c, c->hullFaces, s->hullFaceIndices, s->hullFaceIndexCount);
    upload(c, c->children, s->compoundChildren, s->compoundChildCount);
    c->vertexCount = s->vertexCount; c->meshCount = s->meshCount; c->nodeCount = s->bvhNodeCount;
    c->soup.vertices = c->vertices.p; c->soup.indices = c->indices.p; c->soup.meshes = c->meshes.p;
    c->soup.nodes = c->nodes.p; c->soup.triRefs = c->triRefs.p; c->soup.hullPoints = c->hullPts.p;
    c->soup.hullFaceIndices = c->hullFaces.p; c->soup.children = c->children.p;
}

void validateScene(const vx_body* b, uint32_t n, const vx_soup* s) {
    for (uint32_t i = 0; i < n; ++i) {
        const vx_body& x = b[i];
        if (x.shape > VX_SHAPE_ELLIPSOID) fail(VXHIP_ERR_INVALID, "vxhip: body has an unknown shape");
        if (x.shape == VX_SHAPE_MESH && (!s || x.meshIndex >= s->meshCount))
            fail(VXHIP_ERR_INVALID, "vxhip: mesh body references an invalid mesh");
        if (x.shape == VX_SHAPE_HULL && (!s || x.hullCount < 4 || (uint64_t)x.hullFirst + x.hullCount > s->hullPointCount ||
                                         (uint64_t)x.hullFaceFirst + 3ull * x.hullFaceCount > s->hullFaceIndexCount))
            fail(VXHIP_ERR_INVALID, "vxhip: hull body references invalid point storage");
        if (x.shape == VX_SHAPE_COMPOUND &&
            (!s || x.compoundCount == 0 || (uint64_t)x.compoundFirst + x.compoundCount > s->compoundChildCount))
            fail(VXHIP_ERR_INVALID, "vxhip: compound body references invalid child storage");
    }
}

// The per-call (Backend-equivalent) entry points share the context's device
// buffers with the resident world: they detach it first, so a later
// vxhip_world_step refuses to run on bodies, caches and settings that belong
// to another call sequence (the world is loaded again to resume).
void detachWorld(vxhip_ctx* c) {
    c->loaded = false;
    c->prevCount = 0;
    c->prevExact = false;
    c->evPrevCount = 0;
    c->eventCount = 0;
    c->sleepOn = false;
}

void setBodies(vxhip_ctx* c, const vx_body* b, uint32_t n) {
    if (n != c->n || !c->slStability.p) {  // contact stability counters start at zero (sleep.cpp:64)
        c->slStability.reserve(n ? n : 1);
        HIPCK(hipMemsetAsync(c->slStability.p, 0, (size_t)(n ? n : 1) * sizeof(uint16_t), c->stream));
    }
    c->n = n;
    upload(c, c->bodies, b, n);
    c->anyAsleep = false;
    c->anyHigh = false;
    c->hostDynamic.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        c->hostDynamic[i] = b[i].motionType == VX_MOTION_DYNAMIC;
        c->anyAsleep |= b[i].asleep == 1;
        c->anyHigh |= b[i].motionType == VX_MOTION_DYNAMIC && b[i].ccdTuning.quality == VX_QUALITY_HIGH &&
                      b[i].ccdTuning.enableContinuous;
    }
}

int bodyBits(uint32_t n) { return vxhost::bitsFor(n ? n - 1 : 0) ? vxhost::bitsFor(n ? n - 1 : 0) : 1; }

// --------------------------------------------------------------- stages --
// Broad phase: fills c->pairs (sorted packed contact keys), returns count.
constexpr uint64_t kCellTableMax = 1ull << 22;
// Colours with at most this many contacts are solved in runs by one
// workgroup instead of a launch each.
constexpr uint32_t kNarrowColor = 2048;
// CUs the persistent island-batch grid leaves to the big-island colour chain
// beside it (the chain's LDS tail needs a CU of its own).
constexpr uint32_t kChainCus = 4;
constexpr uint32_t kSleepMaxBodies = 256;
// Joint sets this large go to the device; smaller ones stay with the host
// world, as backend_cuda.cu:696-701 decides.
constexpr uint32_t kMinDeviceJoints = 64;

uint32_t broadPhase(vxhip_ctx* c, float dt) {
    const uint32_t n = c->n;
    hipStream_t s = c->stream;
    c->lo.reserve(n); c->hi.reserve(n);
    c->keys.reserve(n); c->keysTmp.reserve(n); c->idx.reserve(n); c->idxTmp.reserve(n); c->oversize.reserve(n);
    uint32_t* k = c->counters.p;
    kResetCounters<<<1, 128, 0, s>>>(k);
    kAabb<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, c->soup, dt, c->lo.p, c->hi.p, k + cExtent, c->extHist.p);
    kCellSize<<<1, kExtBuckets, 0, s>>>(c->extHist.p, n, k + cExtent);
    kCellBounds<<<blocks(n), kThreads, 0, s>>>(c->lo.p, c->hi.p, n, k + cExtent, (int*)(k + cBounds),
                                               k + cOversize2);
    HIPCK(hipGetLastError());
    readCounters(c);
    const uint32_t* h = c->hostCounters;
    float ext;
    std::memcpy(&ext, &h[cExtent], 4);
    Grid g{};
    g.cell = ext > 0.0f ? ext * 1.0001f : 1.0f;
    const int* bnd = (const int*)(h + cBounds);
    const uint32_t oversizeCount = h[cOversize2];
    if (oversizeCount < n) {
        g.minx = bnd[0]; g.miny = bnd[1]; g.minz = bnd[2];
        g.nx = (uint32_t)((int64_t)bnd[3] - bnd[0] + 1);
        g.ny = (uint32_t)((int64_t)bnd[4] - bnd[1] + 1);
        g.nz = (uint32_t)((int64_t)bnd[5] - bnd[2] + 1);
    } else {
        g.nx = g.ny = g.nz = 1;
    }
    g.ncells = (uint64_t)g.nx * g.ny * g.nz;
    const bool batched = c->worldCount > 1 && c->worldN == n;
    const uint32_t W = batched ? c->worldCount : 1u;
    g.sentinel = (uint64_t)W * g.ncells;
    c->stats.cellSize = g.cell;
    c->stats.oversize = oversizeCount;
    if (batched) {
        c->worldOversize.reserve(W + 1); c->worldStart.reserve(W + 1); c->worldFill.reserve(W);
        c->oversizeByWorld.reserve(n);
        fill32(c, c->worldOversize.p, W, 0);
    }
    kCellKeys<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, c->lo.p, c->hi.p, n, g, batched ? c->worldOf.p : nullptr,
                                             c->keys.p, c->idx.p, c->oversize.p, k + cOversize,
                                             batched ? c->worldOversize.p : nullptr);
    HIPCK(vxhost::radixSort(c->keys.p, c->idx.p, c->keysTmp.p, c->idxTmp.p, n, vxhost::bitsFor(g.sentinel),
                            tempFor(c, vxhost::radixTempBytes(n)), s));
    if (batched && oversizeCount) {
        HIPCK(vxhost::exclusiveScan(c->worldOversize.p, c->worldStart.p, W, c->worldStart.p + W,
                                    tempFor(c, vxhost::scanTempBytes(W)), s));
        fill32(c, c->worldFill.p, W, 0);
        kOversizeByWorld<<<blocks(oversizeCount), kThreads, 0, s>>>(c->oversize.p, oversizeCount, c->worldOf.p,
                                                                    c->worldStart.p, c->worldFill.p,
                                                                    c->oversizeByWorld.p);
    }
    // Cell table for grids of up to kCellTableMax cells; larger grids (many
    // batched worlds) use binary searches.
    const bool table = g.sentinel <= kCellTableMax;
    if (table) {
        c->cellRange.reserve(g.sentinel);
        fill32(c, (uint32_t*)c->cellRange.p, (uint32_t)(2 * g.sentinel), 0);
        kCellRanges<<<blocks(n), kThreads, 0, s>>>(c->keys.p, n, g.sentinel, c->cellRange.p);
    }
    const int nb = bodyBits(n);
    if (c->pairCap == 0) c->pairCap = 32 * n + 1024;
    for (;;) {
        c->pairs.reserve(c->pairCap);
        kResetCounters<<<1, 128, 0, s>>>(k);  // only cPairs matters here; others re-derived
        solveLaunchBegin(c);
        kGridPairs<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, c->lo.p, c->hi.p, c->keys.p, c->idx.p,
                                                  table ? c->cellRange.p : nullptr, n, g, nb, c->pairs.p, k + cPairs,
                                                  c->pairCap);
        solveLaunchEnd(c, VXHIP_KERNEL_GRID_PAIRS, n);
        if (oversizeCount && batched)
            kOversizePairsWorld<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, c->lo.p, c->hi.p, c->worldOf.p,
                                                               c->worldStart.p, c->oversizeByWorld.p, n, g, nb,
                                                               c->pairs.p, k + cPairs, c->pairCap);
        else if (oversizeCount)
            kOversizePairs<<<blocks((uint64_t)oversizeCount * n), kThreads, 0, s>>>(
                c->bodies.p, c->lo.p, c->hi.p, c->oversize.p, oversizeCount, n, g, nb, c->pairs.p, k + cPairs, c->pairCap);
        HIPCK(hipGetLastError());
        readCounters(c);
        uint32_t np = c->hostCounters[cPairs];
        if (np <= c->pairCap) {
            c->pairsCount = np;
            break;
        }
        c->pairCap = np + np / 4;
    }
    const uint32_t np = c->pairsCount;
    c->pairVals.reserve(np ? np : 1);
    c->keysTmp.reserve(np ? np : 1);
    c->idxTmp.reserve(np ? np : 1);
    HIPCK(vxhost::radixSort(c->pairs.p, c->pairVals.p, c->keysTmp.p, c->idxTmp.p, np, 2 * nb,
                            tempFor(c, vxhost::radixTempBytes(np)), s));
    return np;
}


// Narrow phase over the sorted pairs: contacts in canonical order into
// c->contacts, returns the count.
uint32_t narrowPhase(vxhip_ctx* c, uint32_t np, float dt) {
    hipStream_t s = c->stream;
    const int nb = bodyBits(c->n);
    c->slots.reserve((size_t)(np ? np : 1) * kMaxPairContacts);
    c->counts.reserve(np + 1);
    c->offsets.reserve(np + 1);
    if (np) {
        c->pairClass.reserve(np); c->pairClassTmp.reserve(np); c->pairOrder.reserve(np); c->pairOrderTmp.reserve(np);
        kPairClass<<<blocks(np), kThreads, 0, s>>>(c->bodies.p, c->pairs.p, np, nb, c->pairClass.p, c->pairOrder.p);
        HIPCK(vxhost::radixSort(c->pairClass.p, c->pairOrder.p, c->pairClassTmp.p, c->pairOrderTmp.p, np,
                                kPairClassBits, tempFor(c, vxhost::radixTempBytes(np)), s));
        solveLaunchBegin(c);
        kNarrow<<<blocks(np, kCollideThreads), kCollideThreads, 0, s>>>(c->bodies.p, c->soup, dt, c->pairs.p, np, nb, c->pairOrder.p,
                                                c->slots.p, c->counts.p);
        solveLaunchEnd(c, VXHIP_KERNEL_NARROW, np);
        HIPCK(hipGetLastError());
    }
    HIPCK(vxhost::exclusiveScan(c->counts.p, c->offsets.p, np, c->counters.p + cContacts,
                                tempFor(c, vxhost::scanTempBytes(np)), s));
    readCounters(c);
    const uint32_t m = c->hostCounters[cContacts];
    c->contacts.reserve(m ? m : 1);
    if (np) kCompact<<<blocks(np), kThreads, 0, s>>>(c->slots.p, c->counts.p, c->offsets.p, np, c->contacts.p);
    HIPCK(hipGetLastError());
    c->contactCount = m;
    return m;
}

// Greedy coloring of c->contacts[0..m) in order; colours into c->color, returns K.
// Contact sets up to this size are coloured by one workgroup (workgroup-scope
// hand-offs); larger ones across the device.
constexpr uint32_t kColorSingleLimit = 1u << 14;

uint32_t colorContacts(vxhip_ctx* c, const vx_contact* cs, uint32_t m) {
    hipStream_t s = c->stream;
    uint32_t* k = c->counters.p;
    c->color.reserve(m ? m : 1);
    c->colorCount.reserve(m + 65);
    if (m == 0) return 0;
    const int cb = vxhost::bitsFor(m - 1) ? vxhost::bitsFor(m - 1) : 1;
    const int nb = bodyBits(c->n);
    c->keys.reserve(2 * (size_t)m); c->keysTmp.reserve(2 * (size_t)m);
    c->colorVals.reserve(2 * (size_t)m); c->colorVals2.reserve(2 * (size_t)m);
    c->pred.reserve(2 * (size_t)m);
    c->maskAfter.reserve(2 * (size_t)m);
    c->nextAfter.reserve(2 * (size_t)m);
    kResetCounters<<<1, 128, 0, s>>>(k);
    fillMany(c, {{c->pred.p, 2 * m, kNone}, {c->color.p, m, kNone}, {c->colorCount.p, m + 65, 0}});
    kColorEntries<<<blocks(m), kThreads, 0, s>>>(c->bodies.p, cs, m, cb, c->keys.p, c->colorVals.p, k + cEntries);
    HIPCK(hipGetLastError());
    // At most two entries per contact; the count stays on the device.
    HIPCK(vxhost::radixSort(c->keys.p, c->colorVals.p, c->keysTmp.p, c->colorVals2.p, 2 * m, nb + cb,
                            tempFor(c, vxhost::radixTempBytes(2 * m)), s, k + cEntries));
    kColorPreds<<<blocks(2 * m), kThreads, 0, s>>>(c->keys.p, c->colorVals.p, 2 * m, cb, c->pred.p, k + cEntries);
    const bool batched = c->worldCount > 1 && c->worldN == c->n && c->worldsContiguous;
    if (m <= kColorSingleLimit) {
        kGreedyColor<<<1, kColorThreads, 0, s>>>(c->bodies.p, cs, m, c->pred.p, c->maskAfter.p, c->nextAfter.p,
                                                 c->color.p, c->colorCount.p, k + cColors, k + cColorError);
    } else if (batched) {
        // One workgroup per world (the worlds' contact ranges are contiguous).
        const uint32_t W = c->worldCount;
        c->colorSeg.reserve(W + 1); c->colorSegCount.reserve(W);
        fill32(c, c->colorSegCount.p, W, 0);
        kContactWorlds<<<blocks(m), kThreads, 0, s>>>(cs, m, c->worldOf.p, c->colorSegCount.p);
        HIPCK(vxhost::exclusiveScan(c->colorSegCount.p, c->colorSeg.p, W, c->colorSeg.p + W,
                                    tempFor(c, vxhost::scanTempBytes(W)), s));
        kGreedyColorWorlds<<<W, kThreads, 0, s>>>(c->bodies.p, cs, c->colorSeg.p, c->pred.p, c->maskAfter.p,
                                                       c->nextAfter.p, c->color.p, c->colorCount.p, k + cColors,
                                                       k + cColorError);
    } else {
        ++c->stats.launches[VXHIP_LAUNCH_GREEDY_DEVICE];
        kGreedyColorDevice<<<std::min<uint32_t>(blocks(m), kColorBlocks), kThreads, 0, s>>>(
            c->bodies.p, cs, m, c->pred.p, c->maskAfter.p, c->nextAfter.p, c->color.p, c->colorCount.p, k + cColors,
            k + cColorError);
    }
    HIPCK(hipGetLastError());
    readCounters(c);
    if (c->hostCounters[cColorError]) fail(VXHIP_ERR_DEVICE_LOST, "vxhip: coloring did not converge");
    return c->hostCounters[cColors];
}

// Stable bucket of the coloured contacts by colour (backend_cuda.cu:634-640):
// c->order[i] = canonical index of the i-th contact in colour order; colour
// ranges into c->colorStart (device and host).
void colorOrder(vxhip_ctx* c, uint32_t m, uint32_t K) {
    hipStream_t s = c->stream;
    c->order.reserve(m ? m : 1);
    c->colorStart.reserve(K + 2);
    if (m == 0) { c->colorStartHost.assign(1, 0); return; }
    c->keys.reserve(m); c->keysTmp.reserve(m); c->colorVals2.reserve(m);
    kColorKeys<<<blocks(m), kThreads, 0, s>>>(c->color.p, m, c->keys.p, c->order.p);
    HIPCK(vxhost::radixSort(c->keys.p, c->order.p, c->keysTmp.p, c->colorVals2.p, m,
                            vxhost::bitsFor(K ? K - 1 : 0), tempFor(c, vxhost::radixTempBytes(m)), s));
    HIPCK(vxhost::exclusiveScan(c->colorCount.p, c->colorStart.p, K, c->colorStart.p + K,
                                tempFor(c, vxhost::scanTempBytes(K)), s));
    // Read on the host only by the positional correction, after the island
    // set-up's counter readback has synchronised the stream.
    c->colorStartHost.resize(K + 1);
    HIPCK(hipMemcpyAsync(c->colorStartHost.data(), c->colorStart.p, (K + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
}

uint32_t maxColorSize(const std::vector<uint32_t>& start) {
    uint32_t mx = 0;
    for (size_t i = 0; i + 1 < start.size(); ++i) mx = std::max(mx, start[i + 1] - start[i]);
    return mx;
}


// Island partition of the colour-ordered contacts (needs this step's row
// flags): small islands are solved whole by one workgroup each, big ones by
// per-colour launches over the big-island list.
void prepareIslands(vxhip_ctx* c, uint32_t m, uint32_t K, const uint4* meta) {
    hipStream_t s = c->stream;
    uint32_t* k = c->counters.p;
    const uint32_t n = c->n;
    const uint32_t globalBase = m / kBatch + 1;  // above every LDS batch id
    c->uf.reserve(n); c->islandSize.reserve(n);
    c->ldsSize.reserve(n); c->ldsOffset.reserve(n); c->globalFlag.reserve(n); c->globalRank.reserve(n);
    c->bodyRoot.reserve(n); c->bodyRank.reserve(n); c->bodyCount.reserve(n); c->bodyOffset.reserve(n);
    c->bodyList.reserve(n); c->slotOf.reserve(n); c->rangeStart.reserve(globalBase); c->rangeEnd.reserve(globalBase);
    c->islandOf.reserve(m); c->bigFlag.reserve(m); c->bigPos.reserve(m + 1); c->bigIdx.reserve(m);
    c->bigColorCount.reserve(K + 1); c->bigColorStart.reserve(K + 2); c->colorSolve.reserve(m);
    c->smallKeys.reserve(m); c->smallKeysTmp.reserve(m); c->smallIdx.reserve(m); c->smallIdxTmp.reserve(m);
    c->runHead.reserve(m); c->islandHead.reserve(m); c->runPos.reserve(m + 1); c->islandPos.reserve(m + 1);
    c->runStart.reserve(m + 1); c->islandFirstRun.reserve(m + 1); c->slotPair.reserve(m); c->batchBodies.reserve(m);
    kResetCounters<<<1, 128, 0, s>>>(k);
    kIota<<<blocks(n), kThreads, 0, s>>>(c->uf.p, n);
    fillMany(c, {{c->islandSize.p, n, 0}, {c->bodyCount.p, n, 0}, {c->rangeStart.p, globalBase, kNone},
                 {c->rangeEnd.p, globalBase, 0}, {c->bigColorCount.p, K + 1, 0}});
    kColorOfSolve<<<blocks(m), kThreads, 0, s>>>(c->color.p, c->order.p, m, c->colorSolve.p);
    kIslandUnion<<<blocks(m), kThreads, 0, s>>>(meta, m, c->uf.p);
    kIslandOf<<<blocks(m), kThreads, 0, s>>>(meta, m, c->uf.p, c->islandOf.p, c->islandSize.p);
    kIslandKinds<<<blocks(n), kThreads, 0, s>>>(c->islandSize.p, n, c->islandLimit, c->ldsSize.p, c->globalFlag.p,
                                                k + cIslands, k + cLargest);
    HIPCK(vxhost::exclusiveScan(c->ldsSize.p, c->ldsOffset.p, n, nullptr, tempFor(c, vxhost::scanTempBytes(n)), s));
    HIPCK(vxhost::exclusiveScan(c->globalFlag.p, c->globalRank.p, n, nullptr, tempFor(c, vxhost::scanTempBytes(n)), s));
    // LDS slots of the batched islands' dynamic bodies.
    kBatchBodyCount<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, c->uf.p, c->ldsSize.p, c->bodyRoot.p,
                                                   c->bodyRank.p, c->bodyCount.p);
    HIPCK(vxhost::exclusiveScan(c->bodyCount.p, c->bodyOffset.p, n, nullptr, tempFor(c, vxhost::scanTempBytes(n)), s));
    kBatchBodyRange<<<blocks(n), kThreads, 0, s>>>(c->bodyCount.p, c->bodyOffset.p, c->ldsOffset.p, n,
                                                   c->rangeStart.p, c->rangeEnd.p);
    kBatchBodySlots<<<blocks(n), kThreads, 0, s>>>(c->bodyRoot.p, c->bodyRank.p, c->bodyOffset.p, c->ldsOffset.p,
                                                   c->rangeStart.p, n, c->bodyList.p, c->slotOf.p);
    const int mb = vxhost::bitsFor(m - 1) ? vxhost::bitsFor(m - 1) : 1;
    kIslandClassify<<<blocks(m), kThreads, 0, s>>>(c->islandOf.p, c->islandSize.p, c->ldsOffset.p, c->globalRank.p,
                                                   globalBase, m, c->islandLimit, c->bigFlag.p, c->smallKeys.p,
                                                   c->smallIdx.p, k + cSmall, mb);
    HIPCK(vxhost::exclusiveScan(c->bigFlag.p, c->bigPos.p, m, k + cBig, tempFor(c, vxhost::scanTempBytes(m)), s));
    kBigList<<<blocks(m), kThreads, 0, s>>>(c->bigFlag.p, c->bigPos.p, m, c->colorSolve.p, c->bigIdx.p,
                                            c->bigColorCount.p, K);
    HIPCK(vxhost::exclusiveScan(c->bigColorCount.p, c->bigColorStart.p, K, c->bigColorStart.p + K,
                                tempFor(c, vxhost::scanTempBytes(K)), s));
    HIPCK(hipGetLastError());
    readCounters(c);
    const uint32_t ms = c->hostCounters[cSmall];
    c->smallCount = ms;
    c->bigCount = c->hostCounters[cBig];
    c->islandCount = c->hostCounters[cIslands];
    c->bigColorStartHost.resize(K + 1);
    HIPCK(hipMemcpyAsync(c->bigColorStartHost.data(), c->bigColorStart.p, (K + 1) * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, s));
    c->batchCount = c->ldsBatchCount = 0;
    if (ms) {
        const int bb = vxhost::bitsFor(globalBase + m);
        HIPCK(vxhost::radixSort(c->smallKeys.p, c->smallIdx.p, c->smallKeysTmp.p, c->smallIdxTmp.p, ms, bb + mb,
                                tempFor(c, vxhost::radixTempBytes(ms)), s));
        kBatchContactSlots<<<blocks(ms), kThreads, 0, s>>>(c->smallKeys.p, c->smallIdx.p, ms, meta, c->slotOf.p,
                                                           globalBase, mb, c->slotPair.p);
        kRunHeads<<<blocks(ms), kThreads, 0, s>>>(c->smallKeys.p, c->smallIdx.p, ms, c->colorSolve.p, c->runHead.p,
                                                  c->islandHead.p, mb, globalBase, k + cLdsBatches);
        HIPCK(vxhost::exclusiveScan(c->runHead.p, c->runPos.p, ms, k + cRuns, tempFor(c, vxhost::scanTempBytes(ms)), s));
        HIPCK(vxhost::exclusiveScan(c->islandHead.p, c->islandPos.p, ms, k + cBatches,
                                    tempFor(c, vxhost::scanTempBytes(ms)), s));
        kRunTables<<<blocks(ms), kThreads, 0, s>>>(c->smallKeys.p, c->runHead.p, c->islandHead.p, c->runPos.p,
                                                   c->islandPos.p, ms, mb, globalBase, c->rangeStart.p, c->rangeEnd.p,
                                                   c->runStart.p, c->islandFirstRun.p, c->batchBodies.p);
        kRunSentinels<<<1, 1, 0, s>>>(c->runStart.p, k + cRuns, ms, c->islandFirstRun.p, k + cBatches,
                                      k + cLdsBatches, k + cRanges, k + cLdsContacts);
        kBatchPasses<<<blocks(ms), kThreads, 0, s>>>(c->runStart.p, c->islandFirstRun.p, k + cLdsBatches, k + cPasses,
                                                     c->batchBodies.p, k + cBatchMaxBodies, k + cBatchMaxRows);
        HIPCK(hipGetLastError());
    }
    kSolvePositions<<<blocks(m), kThreads, 0, s>>>(c->smallIdx.p, ms, c->bigIdx.p, c->bigCount, c->solvePos.p);
    HIPCK(hipGetLastError());
    readCounters(c);
    c->largestIsland = c->hostCounters[cLargest];
    c->maxBatchPasses = ms ? c->hostCounters[cPasses] : 0;
    c->ldsContactCount = 0;
    c->batchSmall = false;
    if (ms) {
        c->batchCount = c->hostCounters[cBatches];
        c->ldsBatchCount = c->hostCounters[cLdsBatches];
        c->ldsContactCount = c->hostCounters[cLdsContacts];
        // Every LDS batch fits the small layout: three workgroups share a CU.
        c->batchSmall = c->hostCounters[cBatchMaxBodies] <= kSmallBodies && c->hostCounters[cBatchMaxRows] <= kSmallRows;
    }
    // Longest-first hand-out order of the LDS batches.
    c->batchOrderReady = false;
    const uint32_t lb = c->ldsBatchCount;
    if (c->longestFirst && lb > 1) {
        c->batchKeys.reserve(lb); c->batchKeysTmp.reserve(lb); c->batchOrder.reserve(lb); c->batchOrderTmp.reserve(lb);
        kBatchCostKeys<<<blocks(lb), kThreads, 0, s>>>(c->runStart.p, c->islandFirstRun.p, k + cLdsBatches,
                                                       c->batchThreads / 2, c->batchKeys.p, c->batchOrder.p);
        HIPCK(vxhost::radixSort(c->batchKeys.p, c->batchOrder.p, c->batchKeysTmp.p, c->batchOrderTmp.p, lb, 16,
                                tempFor(c, vxhost::radixTempBytes(lb)), s));
        HIPCK(hipGetLastError());
        c->batchOrderReady = true;
    }
}

// The LDS tail of the big-island colours: the last colours that each hold at
// most colorLaunchRows contacts, kTailRows rows in all, their written
// bodies given LDS slots.
void prepareTail(vxhip_ctx* c, uint32_t K, SolveArrays sa) {
    c->tailColor = K;
    if (!c->bigCount) return;
    const std::vector<uint32_t>& cs = c->bigColorStartHost;
    uint32_t t = K, tailRows = 0;
    while (t > 0) {
        const uint32_t cnt = cs[t] - cs[t - 1];
        if (cnt > c->colorLaunchRows || tailRows + cnt > kTailRows || K - t >= kTailColors) break;
        tailRows += cnt;
        --t;
    }
    if (t >= K || cs[K] == cs[t]) return;
    hipStream_t s = c->stream;
    const uint32_t rows = cs[K] - cs[t], n2 = 2 * rows, r0 = c->smallCount + cs[t];
    const int nb = bodyBits(c->n);
    const uint64_t sentinel = 1ull << nb;
    c->tailKeys.reserve(n2); c->tailKeysTmp.reserve(n2); c->tailVals.reserve(n2); c->tailValsTmp.reserve(n2);
    c->tailHead.reserve(n2); c->tailPos.reserve(n2); c->tailRowSlot.reserve(n2); c->tailSlotBody.reserve(n2);
    kTailEntries<<<blocks(rows), kThreads, 0, s>>>(sa, r0, rows, sentinel, c->tailKeys.p, c->tailVals.p);
    HIPCK(vxhost::radixSort(c->tailKeys.p, c->tailVals.p, c->tailKeysTmp.p, c->tailValsTmp.p, n2, nb + 1,
                            tempFor(c, vxhost::radixTempBytes(n2)), s));
    kTailHeads<<<blocks(n2), kThreads, 0, s>>>(c->tailKeys.p, n2, sentinel, c->tailHead.p);
    HIPCK(vxhost::exclusiveScan(c->tailHead.p, c->tailPos.p, n2, c->counters.p + cTailSlots,
                                tempFor(c, vxhost::scanTempBytes(n2)), s));
    kTailSlots<<<blocks(n2), kThreads, 0, s>>>(c->tailKeys.p, c->tailVals.p, n2, sentinel, c->tailHead.p, c->tailPos.p,
                                               reinterpret_cast<uint16_t*>(c->tailRowSlot.p), c->tailSlotBody.p);
    HIPCK(hipGetLastError());
    c->tailColor = t;
}

// The side stream of the current device: one per device for the whole
// process (every context forks onto it), so extra contexts take no extra
// hardware queue.  Default priority; created on first use, never destroyed.
constexpr int kMaxSideDevices = 64;
hipStream_t gSide[kMaxSideDevices] = {};
std::mutex gSideMutex;

hipStream_t sideStream(vxhip_ctx* c) {
    if (c->device < 0 || c->device >= kMaxSideDevices) return c->stream;
    std::lock_guard<std::mutex> lock(gSideMutex);
    if (!gSide[c->device]) HIPCK(hipStreamCreateWithFlags(&gSide[c->device], hipStreamNonBlocking));
    if (!c->forkEv) HIPCK(hipEventCreateWithFlags(&c->forkEv, hipEventDisableTiming));
    if (!c->joinEv) HIPCK(hipEventCreateWithFlags(&c->joinEv, hipEventDisableTiming));
    return gSide[c->device];
}

// One pass (warm start, or 2*I sweeps) over every contact: small islands in
// one launch, big islands colour by colour.  When both exist the island
// batches go to the side stream and run beside the big-island chain: they
// touch no big-island body and no big-island contact.  (Not while
// profiling: the per-launch events time one kernel at a time.)
template <int kMode>
void solvePass(vxhip_ctx* c, SolveArrays sa, uint32_t K, int sweeps) {
    hipStream_t s = c->stream;
    // Independent groups go side by side: the island batches on the side
    // stream, and on the library's stream the big-island chain or, without
    // one, the medium islands (they share no body with the batches).
    const uint32_t globalIslands = c->batchCount - c->ldsBatchCount;
    const bool fork = c->sideOn && c->batchCount && !c->profile &&
                      (c->bigCount || (c->ldsBatchCount && globalIslands));
    const hipStream_t bs = fork ? sideStream(c) : s;
    const hipStream_t is = fork && !c->bigCount ? s : bs;  // medium islands
    if (fork) {
        HIPCK(hipEventRecord(c->forkEv, s));
        HIPCK(hipStreamWaitEvent(bs, c->forkEv, 0));
    }
    if (c->batchCount) {
        const uint32_t lds = c->ldsBatchCount, global = c->batchCount - c->ldsBatchCount;
        if (lds) {
            // Contacts visited: every contact of the LDS batches, once per sweep.
            const uint64_t visits = (uint64_t)c->ldsContactCount * sweeps;
            solveLaunchBegin(c);
            ++c->stats.launches[VXHIP_LAUNCH_BATCH];
            // Persistent: as many workgroups as the CUs hold at once (one with
            // the full LDS layout, two or three with the small one), a few CUs
            // left free for the big-island chain beside them.
            const bool small = c->batchSmall;
            const uint32_t perCu = small ? (c->batchThreads == 512 ? 2u : 3u) : 1u;
            const uint32_t cus = c->bigCount ? std::max<uint32_t>(1u, c->cuCount - kChainCus) : c->cuCount;
            const uint32_t cap = c->persistentBatches ? cus * perCu : 65535u;
            const uint32_t grid = std::min<uint32_t>(lds, cap);
            const uint32_t* order = c->batchOrderReady ? c->batchOrder.p : nullptr;
#define VX_BATCH_LAUNCH(T, SMALL)                                                                                   \
    kBatchSweeps<kMode, T, SMALL><<<grid, T, 0, bs>>>(sa, c->slotPair.p, c->runStart.p, c->islandFirstRun.p,        \
                                                      c->batchBodies.p, c->bodyList.p, c->counters.p + cRanges, order, \
                                                      c->batchQueue.p, sweeps, c->counters.p + cBatchError)
            if (c->batchThreads == 512) {
                if (small) VX_BATCH_LAUNCH(512, true);
                else VX_BATCH_LAUNCH(512, false);
            } else {
                if (small) VX_BATCH_LAUNCH(256, true);
                else VX_BATCH_LAUNCH(256, false);
            }
#undef VX_BATCH_LAUNCH
            solveLaunchEnd(c, kMode == 2 ? VXHIP_KERNEL_BATCH_WARM : VXHIP_KERNEL_BATCH, visits);
        }
        if (global) {
            solveLaunchBegin(c);
            ++c->stats.launches[VXHIP_LAUNCH_ISLAND];
            kIslandSweeps<kMode, kThreads><<<std::min<uint32_t>(global, 4096), kThreads, 0, is>>>(
                sa, c->runStart.p, c->islandFirstRun.p, c->counters.p + cRanges + 2, sweeps);
            solveLaunchEnd(c, VXHIP_KERNEL_ISLAND, (uint64_t)(c->smallCount - c->ldsContactCount) * sweeps);
        }
    }
    if (c->bigCount) {
        const std::vector<uint32_t>& cs = c->bigColorStartHost;
        for (int it = 0; it < sweeps; ++it)
            for (uint32_t col = 0; col < K;) {
                const uint32_t cnt = cs[col + 1] - cs[col];
                if (!cnt) { ++col; continue; }
                if (kMode != 2 && col == c->tailColor) {
                    solveLaunchBegin(c);
                    ++c->stats.launches[VXHIP_LAUNCH_TAIL];
                    kSweepTailPair<kMode><<<1, kSweepRunThreads, 0, s>>>(sa, c->smallCount, c->bigColorStart.p, col, K,
                                                                         c->tailRowSlot.p, c->tailSlotBody.p,
                                                                         c->counters.p + cTailSlots);
                    solveLaunchEnd(c, VXHIP_KERNEL_COLOR, cs[K] - cs[col]);
                    break;
                }
                if (cnt > c->colorLaunchRows) {
                    solveLaunchBegin(c);
                    ++c->stats.launches[VXHIP_LAUNCH_COLOR_PAIR];
                    kSweepColorPair<kMode><<<blocks(2 * cnt), kThreads, 0, s>>>(sa, c->smallCount, cs[col], cnt);
                    solveLaunchEnd(c, VXHIP_KERNEL_COLOR, cnt);
                    ++col;
                    continue;
                }
                uint32_t end = col + 1;
                while (end < K && (kMode == 2 || end != c->tailColor) && cs[end + 1] - cs[end] <= c->colorLaunchRows)
                    ++end;
                solveLaunchBegin(c);
                ++c->stats.launches[VXHIP_LAUNCH_COLOR_RUN];
                kSweepColorRun<kMode><<<1, kSweepRunThreads, 0, s>>>(sa, c->smallCount, c->bigColorStart.p, col, end);
                solveLaunchEnd(c, VXHIP_KERNEL_COLOR, cs[end] - cs[col]);
                col = end;
            }
    }
    if (fork) {
        HIPCK(hipEventRecord(c->joinEv, bs));
        HIPCK(hipStreamWaitEvent(s, c->joinEv, 0));
    }
    HIPCK(hipGetLastError());
}

// One substep's velocity solve (backend_cuda.cu:652-689, 730-747): compact
// body arrays, this substep's rows, warm start on the first substep, then
// 2*I colour-ordered sweeps; velocities (and impulses after the last
// substep) written back.
void substepSolve(vxhip_ctx* c, vx_contact* solve, uint32_t K, float h, const vx_solver_options& opt, bool first,
                  bool last) {
    const uint32_t n = c->n, m = c->solveCount;
    hipStream_t s = c->stream;
    c->sVel.reserve(n); c->sAng.reserve(n); c->sInvI.reserve(n); c->sFrame.reserve(n);
    c->sRows.reserve(8 * (size_t)m); c->sImp.reserve(2 * (size_t)m); c->sMeta.reserve(m); c->sMetaK.reserve(m);
    c->sSide.reserve(8 * (size_t)m); c->solvePos.reserve(m);
    SolveArrays sa{c->sVel.p, c->sAng.p, c->sInvI.p, c->sFrame.p, c->sRows.p, c->sMeta.p, c->sImp.p, c->sSide.p,
                   c->solvePos.p};
    kGatherSolveBodies<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, sa);
    if (first) {
        // Islands, batches and solve positions for this step's contact set.
        kPrepareMeta<<<blocks(m), kThreads, 0, s>>>(c->bodies.p, solve, m, c->sMetaK.p);
        prepareIslands(c, m, K, c->sMetaK.p);
        kLoadImpulses<<<blocks(m), kThreads, 0, s>>>(solve, m, c->sImp.p, c->solvePos.p);
    }
    kPrepareRows<<<blocks(m), kThreads, 0, s>>>(c->bodies.p, solve, m, h, sa);
    if (first) prepareTail(c, K, sa);
    if (first) solvePass<2>(c, sa, K, 1);
    const int sweeps = 2 * opt.velocityIterations;
    if (opt.frictionModel == VX_FRICTION_CONE_BLOCK) solvePass<1>(c, sa, K, sweeps);
    else solvePass<0>(c, sa, K, sweeps);
    kScatterVelocities<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, sa);
    if (last) kStoreImpulses<<<blocks(m), kThreads, 0, s>>>(solve, m, c->sImp.p, c->solvePos.p);
    HIPCK(hipGetLastError());
}

void positionCorrection(vxhip_ctx* c, const vx_contact* solve, uint32_t K, float h, const vx_solver_options& opt) {
    const std::vector<uint32_t>& cs = c->colorStartHost;
    for (int it = 0; it < opt.positionIterations; ++it)
        for (uint32_t col = 0; col < K;) {
            const uint32_t cnt = cs[col + 1] - cs[col];
            if (!cnt) { ++col; continue; }
            if (cnt > kNarrowColor) {
                kPositionColor<<<blocks(cnt), kThreads, 0, c->stream>>>(c->bodies.p, solve, c->colorStart.p, col, h, opt);
                ++col;
                continue;
            }
            uint32_t end = col + 1;
            while (end < K && cs[end + 1] - cs[end] <= kNarrowColor) ++end;
            kPositionColorRun<<<1, kRunThreads, 0, c->stream>>>(c->bodies.p, solve, c->colorStart.p, col, end, h, opt);
            col = end;
        }
    HIPCK(hipGetLastError());
}

// prepareJoints (backend_cuda.cu:800-843): greedy colours in joint order
// (first free bit of the two dynamic endpoints' 64-bit masks, else past both
// bodies' last colour), a stable colour order, and the sorted keys of the
// pairs the joints keep from colliding (world.cpp:4045-4066).  Host work on
// the joint list, as in the reference backend; redone when joints break.
void prepareJoints(vxhip_ctx* c) {
    const std::vector<vx_joint>& js = c->hostJoints;
    const uint32_t nj = (uint32_t)js.size();
    c->jointCount = nj;
    c->jointColors = 0;
    c->jointExcludedCount = 0;
    if (!nj) return;
    std::vector<uint64_t> mask(c->n, 0);
    std::vector<int> next(c->n, 64), colorOf(nj);
    int numColors = 0;
    for (uint32_t k = 0; k < nj; ++k) {
        const bool sa = !c->hostDynamic[js[k].a], sb = !c->hostDynamic[js[k].b];
        const uint64_t used = (sa ? 0 : mask[js[k].a]) | (sb ? 0 : mask[js[k].b]);
        int color;
        if (~used) {
            color = 0;
            while (used & (1ull << color)) ++color;
        } else {
            const int na = sa ? 64 : next[js[k].a], nb = sb ? 64 : next[js[k].b];
            color = na > nb ? na : nb;
            if (!sa) next[js[k].a] = color + 1;
            if (!sb) next[js[k].b] = color + 1;
        }
        if (color < 64) {
            if (!sa) mask[js[k].a] |= 1ull << color;
            if (!sb) mask[js[k].b] |= 1ull << color;
        }
        colorOf[k] = color;
        if (color + 1 > numColors) numColors = color + 1;
    }
    // Components of the joint graph over dynamic bodies: joints of different
    // components share no body they write, so each component may run its
    // colours apart from the others, and the per-colour order of the
    // reference holds within it.
    std::vector<uint32_t> parent(c->n);
    for (uint32_t i = 0; i < c->n; ++i) parent[i] = i;
    auto find = [&](uint32_t x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    for (uint32_t k = 0; k < nj; ++k)
        if (c->hostDynamic[js[k].a] && c->hostDynamic[js[k].b]) {
            const uint32_t ra = find(js[k].a), rb = find(js[k].b);
            if (ra != rb) parent[ra < rb ? rb : ra] = ra < rb ? ra : rb;
        }
    // Component of each joint (its first joint's index), components in the
    // order of their first joint.
    std::vector<uint32_t> compOf(nj), compFirst(c->n, kNone);
    std::vector<uint32_t> comps;  // first joint of each component
    for (uint32_t k = 0; k < nj; ++k) {
        const uint32_t body = c->hostDynamic[js[k].a] ? js[k].a : js[k].b;
        if (!c->hostDynamic[body]) { compOf[k] = k; comps.push_back(k); continue; }  // no dynamic body
        const uint32_t r = find(body);
        if (compFirst[r] == kNone) { compFirst[r] = k; comps.push_back(k); }
        compOf[k] = compFirst[r];
    }
    std::vector<uint32_t> compSize(nj, 0), compRank(nj, 0);
    for (uint32_t k = 0; k < nj; ++k) ++compSize[compOf[k]];
    // Packs: whole components in order, closed once they reach kJointPack.
    std::vector<uint32_t> packOf(nj);
    std::vector<uint32_t> packSize;
    {
        std::vector<uint32_t> packOfComp(nj, 0);
        uint32_t cur = 0;
        for (uint32_t f : comps) {
            if (packSize.empty() || cur >= kJointPack) { packSize.push_back(0); cur = 0; }
            packOfComp[f] = (uint32_t)packSize.size() - 1;
            packSize.back() += compSize[f];
            cur += compSize[f];
        }
        for (uint32_t k = 0; k < nj; ++k) packOf[k] = packOfComp[compOf[k]];
    }
    const uint32_t np = (uint32_t)packSize.size();
    // Pack runs (pack, colour) in joint order; oversize packs' joints go to
    // the per-colour list.
    std::vector<std::vector<uint32_t>> runs;  // per (pack, colour): joints
    std::vector<uint32_t> packRunsHost, packOrderHost, packFirst;
    std::vector<uint8_t> bigJoint(nj, 0);
    {
        std::vector<std::vector<uint32_t>> byPack(np);
        for (uint32_t k = 0; k < nj; ++k) byPack[packOf[k]].push_back(k);
        for (uint32_t p = 0; p < np; ++p) {
            if (packSize[p] > c->jointPackMax) {
                for (uint32_t k : byPack[p]) bigJoint[k] = 1;
                continue;
            }
            int colours = 0;
            for (uint32_t k : byPack[p]) colours = colorOf[k] + 1 > colours ? colorOf[k] + 1 : colours;
            std::vector<uint32_t> start(colours + 1, 0);
            for (uint32_t k : byPack[p]) ++start[colorOf[k] + 1];
            for (int q = 0; q < colours; ++q) start[q + 1] += start[q];
            std::vector<uint32_t> fillP(start.begin(), start.end() - 1), ordP(byPack[p].size());
            for (uint32_t k : byPack[p]) ordP[fillP[colorOf[k]]++] = k;
            packFirst.push_back((uint32_t)packRunsHost.size());
            const uint32_t base = (uint32_t)packOrderHost.size();
            for (int q = 0; q < colours; ++q)
                if (start[q + 1] > start[q]) packRunsHost.push_back(base + start[q]);
            packOrderHost.insert(packOrderHost.end(), ordP.begin(), ordP.end());
        }
        packFirst.push_back((uint32_t)packRunsHost.size());
        packRunsHost.push_back((uint32_t)packOrderHost.size());
    }
    c->jointPacks = (uint32_t)packFirst.size() - 1;
    if (c->jointPacks) {
        upload(c, c->packOrder, packOrderHost.data(), (uint32_t)packOrderHost.size());
        upload(c, c->packRunStart, packRunsHost.data(), (uint32_t)packRunsHost.size());
        upload(c, c->packFirstRun, packFirst.data(), (uint32_t)packFirst.size());
    }
    // The per-colour list holds the oversize packs' joints only.
    c->jointColors = (uint32_t)numColors;
    c->jointColorStartHost.assign(numColors + 1, 0);
    for (uint32_t k = 0; k < nj; ++k)
        if (bigJoint[k]) ++c->jointColorStartHost[colorOf[k] + 1];
    for (int k = 0; k < numColors; ++k) c->jointColorStartHost[k + 1] += c->jointColorStartHost[k];
    std::vector<uint32_t> fill(c->jointColorStartHost.begin(), c->jointColorStartHost.end() - 1), order(nj);
    for (uint32_t k = 0; k < nj; ++k)
        if (bigJoint[k]) order[fill[colorOf[k]]++] = k;
    std::vector<uint64_t> excluded;
    for (const vx_joint& j : js)
        if (!j.collideConnected) {
            const uint32_t lo = j.a < j.b ? j.a : j.b, hi = j.a < j.b ? j.b : j.a;
            excluded.push_back((uint64_t)lo << 32 | hi);
        }
    std::sort(excluded.begin(), excluded.end());
    c->jointExcludedCount = (uint32_t)excluded.size();
    upload(c, c->joints, js.data(), nj);
    upload(c, c->jointOrder, order.data(), nj);
    if (!excluded.empty()) upload(c, c->jointExcluded, excluded.data(), (uint32_t)excluded.size());
    HIPCK(hipStreamSynchronize(c->stream));
}

// Drops the contacts of joint-connected pairs (order kept) and takes their
// pairs out of the CCD net; returns the new contact count.
uint32_t jointContactsOff(vxhip_ctx* c, uint32_t np, uint32_t m) {
    if (!c->jointExcludedCount) return m;
    hipStream_t s = c->stream;
    const int nb = bodyBits(c->n);
    if (np) kJointPairsOff<<<blocks(np), kThreads, 0, s>>>(c->pairs.p, np, nb, c->jointExcluded.p,
                                                            c->jointExcludedCount, c->counts.p);
    if (!m) return m;
    c->jointKeep.reserve(m); c->jointKeepPos.reserve(m); c->contactsTmp.reserve(m);
    kJointContactsOff<<<blocks(m), kThreads, 0, s>>>(c->contacts.p, m, c->jointExcluded.p, c->jointExcludedCount,
                                                      c->jointKeep.p);
    HIPCK(vxhost::exclusiveScan(c->jointKeep.p, c->jointKeepPos.p, m, c->counters.p + cContacts,
                                tempFor(c, vxhost::scanTempBytes(m)), s));
    kJointContactsCompact<<<blocks(m), kThreads, 0, s>>>(c->contacts.p, m, c->jointKeep.p, c->jointKeepPos.p,
                                                          c->contactsTmp.p);
    HIPCK(hipGetLastError());
    readCounters(c);
    std::swap(c->contacts.p, c->contactsTmp.p);
    std::swap(c->contacts.cap, c->contactsTmp.cap);
    c->contactCount = c->hostCounters[cContacts];
    return c->contactCount;
}

// The joints of one substep (backend_cuda.cu:738-745): reset, 8 passes over
// the colours, break marking.
void substepJoints(vxhip_ctx* c, float h, bool first) {
    const uint32_t nj = c->jointCount;
    if (!nj) return;
    hipStream_t s = c->stream;
    kJointReset<<<blocks(nj), kThreads, 0, s>>>(c->joints.p, nj, first ? 0 : 1);
    if (c->jointPacks)
        kJointPacks<<<c->jointPacks, kJointPackThreads, 0, s>>>(c->bodies.p, c->joints.p, c->packOrder.p,
                                                                c->packRunStart.p, c->packFirstRun.p, h);
    if (!c->jointColorStartHost.empty() && c->jointColorStartHost.back())
        for (int it = 0; it < vxd::kJointIterations; ++it)
        for (uint32_t col = 0; col < c->jointColors; ++col) {
            const uint32_t first0 = c->jointColorStartHost[col], cnt = c->jointColorStartHost[col + 1] - first0;
            if (cnt) kJointColor<<<blocks(cnt), kThreads, 0, s>>>(c->bodies.p, c->joints.p, c->jointOrder.p, first0, cnt, h);
        }
    kJointMarkBroken<<<blocks(nj), kThreads, 0, s>>>(c->joints.p, nj, h);
    HIPCK(hipGetLastError());
}

__global__ void kToiInit(vxk::ToiState* st, float dt) {
    if (threadIdx.x || blockIdx.x) return;
    st->remaining = dt;
    st->bestBits = 0xffffffffu;
    st->bestPair = ~0ull;
    st->prevA = st->prevB = vxk::kNone;
    st->total = st->eligible = st->statics = 0;
}

__global__ void kToiReset(vxk::ToiState* st) {
    if (threadIdx.x || blockIdx.x) return;
    st->bestBits = 0xffffffffu;
    st->bestPair = ~0ull;
}

// processHighToi (world.cpp:3019-3142): the event loop runs on the host, one
// device search per event (a few per step in practice).
void highToi(vxhip_ctx* c, uint32_t m, float dt) {
    c->toiProcessed = false;
    c->toiEventCount = 0;
    const vxk::ToiSettings& ts = c->toiSettings;
    if (!c->anyHigh || !ts.enabled || !ts.splitting) return;
    hipStream_t s = c->stream;
    const uint32_t n = c->n;
    c->toiEligible.reserve(n); c->toiE.reserve(n); c->toiS.reserve(n); c->toiEvents.reserve(n); c->toiState.reserve(1);
    vxk::ToiState* st = c->toiState.p;
    kToiEligible<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, c->toiStarts.p, n, ts, c->toiEligible.p);
    if (m) kToiContactPass<<<1, kWakeThreads, 0, s>>>(c->bodies.p, c->solve.p, m, c->toiEligible.p);
    kToiInit<<<1, 1, 0, s>>>(st, dt);
    fill32(c, c->toiEvents.p, n, 0u);
    kToiRewind<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, c->toiStarts.p, n, c->toiEligible.p, c->toiE.p, c->toiS.p, st);
    HIPCK(hipGetLastError());
    vxk::ToiState h;
    HIPCK(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    const uint32_t ne = h.eligible, ns = h.statics;
    if (!ne) return;
    c->toiProcessed = true;
    const uint64_t P = (uint64_t)ne * (ne + ns);
    c->toiBits.reserve(P); c->toiHit.reserve(2 * P);
    while (h.remaining > 1e-6f && h.total < ts.maxTotal) {
        kToiReset<<<1, 1, 0, s>>>(st);
        kToiPairs<<<blocks(P), kThreads, 0, s>>>(c->bodies.p, c->soup, ts, c->toiE.p, c->toiS.p, c->toiEvents.p, ne, ns,
                                                st, c->toiBits.p, c->toiHit.p);
        kToiPick<<<blocks(P), kThreads, 0, s>>>(c->toiE.p, c->toiS.p, ne, ns, c->toiBits.p, st);
        kToiSelect<<<blocks(P), kThreads, 0, s>>>(c->toiE.p, c->toiS.p, ne, ns, c->toiBits.p, c->toiHit.p, st);
        kToiAdvance<<<blocks(ne), kThreads, 0, s>>>(c->bodies.p, c->toiE.p, ne, st, 0);
        kToiImpulse<<<1, 1, 0, s>>>(c->bodies.p, ts, c->toiEligible.p, c->toiEvents.p, st);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        if (h.bestBits == 0xffffffffu) break;
    }
    kToiAdvance<<<blocks(ne), kThreads, 0, s>>>(c->bodies.p, c->toiE.p, ne, st, 1);
    HIPCK(hipGetLastError());
    c->toiEventCount = h.total;
}

void ccdNet(vxhip_ctx* c, uint32_t np, float dt) {
    if (np == 0) return;
    hipStream_t s = c->stream;
    uint32_t* k = c->counters.p;
    const uint32_t n = c->n;
    const int nb = bodyBits(n);
    const int pbits = vxhost::bitsFor(np - 1) ? vxhost::bitsFor(np - 1) : 1;
    c->ccdFlag.reserve(np);
    c->ccdGap.reserve(np);
    const uint8_t* toiDone = c->toiProcessed ? c->toiEligible.p : nullptr;
    solveLaunchBegin(c);
    kCcdCheck<<<blocks(np, kCollideThreads), kCollideThreads, 0, s>>>(c->bodies.p, c->soup, dt, c->pairs.p, c->counts.p, np, nb,
                                              c->pairOrder.p, c->ccdFlag.p,
                                              c->ccdGap.p, k + cCcdFlagged, k + cCcdFirst, toiDone);
    solveLaunchEnd(c, VXHIP_KERNEL_CCD_CHECK, np);
    HIPCK(hipGetLastError());
    // No host round trip for the flagged count: with none flagged, the closure
    // finds no relevant pair and every later launch of the net returns at once.
    // Pairs that can see a rewound body (flagged pairs and what follows them
    // on shared dynamic bodies).
    c->ccdRel.reserve(np); c->ccdRelPos.reserve(np); c->ccdList.reserve(np); c->ccdFirstRel.reserve(n);
    // Eight closure rounds queued at once: a round after one that added
    // nothing returns at once, and the all-pairs fallback runs only when the
    // eighth still added pairs (no host round trip per round).
    fillMany(c, {{c->ccdFirstRel.p, n, kNone}, {k + cCcdRound, kCcdRounds, 0}});
    kCcdFlagsToRel<<<blocks(np), kThreads, 0, s>>>(c->bodies.p, c->pairs.p, c->ccdFlag.p, np, nb, c->ccdRel.p,
                                                   c->ccdFirstRel.p);
    for (int round = 0; round < kCcdRounds; ++round) {
        const uint32_t* prev = round ? k + cCcdRound + round - 1 : nullptr;
        kCcdRelRound<<<blocks(np), kThreads, 0, s>>>(c->bodies.p, c->pairs.p, c->counts.p, np, nb, c->ccdFirstRel.p,
                                                     c->ccdRel.p, k + cCcdRound + round, prev);
    }
    kCcdRelAll<<<blocks(np), kThreads, 0, s>>>(c->counts.p, np, c->ccdRel.p, k + cCcdRound + kCcdRounds - 1);
    HIPCK(hipGetLastError());
    HIPCK(vxhost::exclusiveScan(c->ccdRel.p, c->ccdRelPos.p, np, k + cCcdRelevant, tempFor(c, vxhost::scanTempBytes(np)), s));
    kCcdRelList<<<blocks(np), kThreads, 0, s>>>(c->ccdRel.p, c->ccdRelPos.p, np, c->ccdList.p);
    HIPCK(hipGetLastError());
    // The relevant pairs' count stays on the device: every buffer is sized
    // for all np pairs and the kernels read the count.
    c->keys.reserve(2 * (size_t)np); c->keysTmp.reserve(2 * (size_t)np);
    c->colorVals.reserve(2 * (size_t)np); c->colorVals2.reserve(2 * (size_t)np);
    c->pred.reserve(2 * (size_t)np);
    c->ccdDone.reserve(np);
    fillMany(c, {{c->pred.p, 2 * np, kNone}, {c->ccdDone.p, np, 0}});
    kCcdEntries<<<blocks(np), kThreads, 0, s>>>(c->bodies.p, c->pairs.p, c->ccdList.p, k + cCcdRelevant, nb, pbits,
                                                c->keys.p, c->colorVals.p, k + cEntries);
    HIPCK(hipGetLastError());
    HIPCK(vxhost::radixSort(c->keys.p, c->colorVals.p, c->keysTmp.p, c->colorVals2.p, 2 * np, nb + pbits,
                            tempFor(c, vxhost::radixTempBytes(2 * np)), s, k + cEntries));
    kColorPreds<<<blocks(2 * np), kThreads, 0, s>>>(c->keys.p, c->colorVals.p, 2 * np, pbits, c->pred.p, k + cEntries);
    // A co-resident grid; entry t goes to wavefront t mod (grid x 4).
    solveLaunchBegin(c);
    const uint32_t ccdGrid = std::max<uint32_t>(1u, c->cuCount * kCcdGridPerCu4 / 4);
    kCcdDataflow<<<ccdGrid, kThreads, 0, s>>>(c->bodies.p, c->prevPos.p, c->prevQ.p, c->soup, dt, c->pairs.p,
                                              c->ccdList.p, k + cCcdRelevant, nb, c->ccdFlag.p, c->ccdGap.p,
                                              c->pred.p, c->ccdDone.p, k + cCcdExamined, k + cCcdRescued,
                                              k + cColorError, toiDone);
    solveLaunchEnd(c, VXHIP_KERNEL_CCD_DATAFLOW, 0);
    HIPCK(hipGetLastError());
}

// Contact events (world.cpp:4399-4485) from the step's solved contacts; the
// counts are read with the step's last counter readback.
void contactEvents(vxhip_ctx* c, uint32_t m, float dt) {
    hipStream_t s = c->stream;
    uint32_t* k = c->counters.p;
    const uint32_t n = c->n, np = c->evPrevCount;
    const int nb = bodyBits(n);
    const uint64_t inactive = 1ull << (2 * nb);
    const size_t keyCap = (size_t)m + np + 1;
    c->evKeys.reserve(m + 1); c->evKeysTmp.reserve(m + 1); c->evVals.reserve(m + 1); c->evValsTmp.reserve(m + 1);
    c->evHead.reserve(m + 1); c->evRunPos.reserve(m + 1);
    c->evEndP.reserve(np + 1); c->evEndBits.reserve(np + 1); c->evApp.reserve(np + 1); c->evEmit.reserve(np + 1);
    c->evPrevKeys.grow(keyCap, np, s);  // holds the previous step's keys
    c->evCurKeys.reserve(keyCap); c->evMerged.reserve(keyCap); c->evAppKeys.reserve(np + 1);
    c->events.reserve(keyCap);
    if (m) {
        kEventKeys<<<blocks(m), kThreads, 0, s>>>(c->bodies.p, c->solve.p, m, dt, nb, inactive, c->evKeys.p, c->evVals.p,
                                                  k + cEvActive);
        HIPCK(vxhost::radixSort(c->evKeys.p, c->evVals.p, c->evKeysTmp.p, c->evValsTmp.p, m, 2 * nb + 1,
                                tempFor(c, vxhost::radixTempBytes(m)), s));
        kEventRunHeads<<<blocks(m), kThreads, 0, s>>>(c->evKeys.p, k + cEvActive, m, c->evHead.p);
    }
    HIPCK(vxhost::exclusiveScan(c->evHead.p, c->evRunPos.p, m, k + cEvRuns, tempFor(c, vxhost::scanTempBytes(m)), s));
    if (m)
        kEventRuns<<<blocks(m), kThreads, 0, s>>>(c->bodies.p, c->solve.p, c->evKeys.p, c->evVals.p, c->evHead.p,
                                                  c->evRunPos.p, k + cEvActive, m, nb, c->evPrevKeys.p, np,
                                                  c->events.p, c->evCurKeys.p);
    if (np) {
        kEventEndsClassify<<<blocks(np), kThreads, 0, s>>>(c->bodies.p, n, nb, c->evPrevKeys.p, np, c->evCurKeys.p,
                                                           k + cEvRuns, c->evEndP.p, c->evEndBits.p, c->evApp.p);
        HIPCK(vxhost::exclusiveScan(c->evApp.p, c->evApp.p, np, k + cEvInert, tempFor(c, vxhost::scanTempBytes(np)), s));
        kEventEndsEmit<<<blocks(np), kThreads, 0, s>>>(c->evPrevKeys.p, np, k + cEvRuns, c->evEndP.p, c->evEndBits.p,
                                                       c->evApp.p, c->evEmit.p, c->evAppKeys.p);
        HIPCK(vxhost::exclusiveScan(c->evEmit.p, c->evEmit.p, np, k + cEvEnds, tempFor(c, vxhost::scanTempBytes(np)), s));
        kEventEndsWrite<<<blocks(np), kThreads, 0, s>>>(c->bodies.p, nb, c->evPrevKeys.p, np, k + cEvRuns, c->evEndP.p,
                                                        c->evEndBits.p, c->evApp.p, c->evEmit.p, c->prev.p,
                                                        c->prevCount, c->events.p);
        kEventKeysMerge<<<blocks(m + np), kThreads, 0, s>>>(c->evCurKeys.p, k + cEvRuns, c->evAppKeys.p, k + cEvInert,
                                                            (uint32_t)(m + np), c->evMerged.p);
    }
    HIPCK(hipGetLastError());
}

// Host side of the events after the step's counter readback: event count and
// the key list handed to the next step.
void finishEvents(vxhip_ctx* c) {
    const uint32_t* h = c->hostCounters;
    const uint32_t runs = h[cEvRuns];
    if (c->evPrevCount) {
        c->eventCount = runs + h[cEvEnds];
        std::swap(c->evPrevKeys.p, c->evMerged.p);
        std::swap(c->evPrevKeys.cap, c->evMerged.cap);
        c->evPrevCount = runs + h[cEvInert];
    } else {
        c->eventCount = runs;
        std::swap(c->evPrevKeys.p, c->evCurKeys.p);
        std::swap(c->evPrevKeys.cap, c->evCurKeys.cap);
        c->evPrevCount = runs;
    }
}

// World::updateSleeping (world.cpp:3788-3794, sleep.cpp:50-325) over the
// step's contacts, after the events.
void sleepUpdate(vxhip_ctx* c, uint32_t m, float dt) {
    hipStream_t s = c->stream;
    c->sleepParams.dt = dt;
    uint32_t* k = c->counters.p;
    const uint32_t n = c->n;
    if (!n) return;
    c->slParent.reserve(n); c->slFlags.reserve(n); c->slIslandMin.reserve(n); c->slRootState.reserve(n);
    kSleepInit<<<blocks(n), kThreads, 0, s>>>(n, c->slParent.p, c->slFlags.p, c->slIslandMin.p, c->slRootState.p);
    if (m)
        kSleepContacts<<<blocks(m), kThreads, 0, s>>>(c->bodies.p, n, c->solve.p, m, c->sleepParams, c->slParent.p,
                                                      c->slFlags.p);
    if (c->jointCount)
        kSleepJoints<<<blocks(c->jointCount), kThreads, 0, s>>>(c->bodies.p, n, c->joints.p, c->jointCount, c->slParent.p);
    kSleepBodies<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, c->sleepParams, c->slParent.p, c->slFlags.p,
                                                c->slStability.p, c->slIslandMin.p, k + cSlWoke);
    kSleepTransitions<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, c->sleepParams, c->slParent.p, c->slIslandMin.p,
                                                     c->slRootState.p, k + cSlDynamic);
    kSleepIslands<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, c->slParent.p, c->slRootState.p, k + cSlIslands);
    HIPCK(hipGetLastError());
}

// Orders `n` (pair key, value) entries, given in the order the reference
// hands them to std::sort, exactly as that std::sort leaves them: one
// std::sort per world (a stable world-major arrangement first when worlds
// are batched), by the parallel restatement of vx_sort.hip.
void exactOrder(vxhip_ctx* c, uint64_t* keys, uint32_t* vals, uint32_t n) {
    if (n <= 1) return;
    hipStream_t s = c->stream;
    const bool batched = c->worldCount > 1 && c->worldN == c->n;
    const uint32_t W = batched ? c->worldCount : 1u;
    const uint32_t* segStart = nullptr;
    if (batched) {
        const int nb = bodyBits(c->n);
        c->exWorld.reserve(n); c->exWorldTmp.reserve(n); c->exIdx.reserve(n); c->exIdxTmp.reserve(n);
        c->exKeys.reserve(n); c->exVals.reserve(n); c->exCount.reserve(W); c->exStart.reserve(W + 1);
        fill32(c, c->exCount.p, W, 0);
        kEntryWorlds<<<blocks(n), kThreads, 0, s>>>(keys, n, nb, c->worldOf.p, c->exWorld.p, c->exIdx.p, c->exCount.p);
        HIPCK(vxhost::radixSort(c->exWorld.p, c->exIdx.p, c->exWorldTmp.p, c->exIdxTmp.p, n,
                                vxhost::bitsFor(W - 1), tempFor(c, vxhost::radixTempBytes(n)), s));
        HIPCK(vxhost::exclusiveScan(c->exCount.p, c->exStart.p, W, c->exStart.p + W,
                                    tempFor(c, vxhost::scanTempBytes(W)), s));
        kPermuteKV<<<blocks(n), kThreads, 0, s>>>(keys, vals, c->exIdx.p, n, c->exKeys.p, c->exVals.p);
        HIPCK(hipMemcpyAsync(keys, c->exKeys.p, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        HIPCK(hipMemcpyAsync(vals, c->exVals.p, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        segStart = c->exStart.p;
    }
    HIPCK(vxhost::stdSort(keys, vals, n, segStart, W, tempFor(c, vxhost::stdSortTempBytes(n, W)), s));
}

// Puts the canonical warm-start cache into the reference's order: the order
// std::sort leaves the previous step's contacts in, taken in solve order.
// prevRank follows its entries (each keeps its solve position).
void exactPrev(vxhip_ctx* c) {
    const uint32_t mp = c->prevCount;
    hipStream_t s = c->stream;
    const int nb = bodyBits(c->n);
    c->cacheKeys.reserve(mp + 1); c->cacheVals.reserve(mp + 1); c->prevNext.reserve(mp + 1);
    c->carryRank.reserve(mp + 1);
    kPrevSolveKeys<<<blocks(mp), kThreads, 0, s>>>(c->prev.p, mp, c->prevRank.p, nb, c->cacheKeys.p, c->cacheVals.p);
    exactOrder(c, c->cacheKeys.p, c->cacheVals.p, mp);
    kGatherBy<<<blocks(mp), kThreads, 0, s>>>(c->prev.p, c->cacheVals.p, mp, c->prevNext.p);
    kGatherU32<<<blocks(mp), kThreads, 0, s>>>(c->prevRank.p, c->cacheVals.p, mp, c->carryRank.p);
    HIPCK(hipGetLastError());
    std::swap(c->prev.p, c->prevNext.p);
    std::swap(c->prev.cap, c->prevNext.cap);
    std::swap(c->prevRank.p, c->carryRank.p);
    std::swap(c->prevRank.cap, c->carryRank.cap);
    c->prevExact = true;
}

// The warm-start cache for the next step (world.cpp:4492-4521): this step's
// contacts plus the first cached contact of every pair kept by an inert
// Persist event, ordered by pair key; in a step where a body fell asleep, in
// exactly the order std::sort leaves them.
void buildCache(vxhip_ctx* c, uint32_t m) {
    hipStream_t s = c->stream;
    uint32_t* k = c->counters.p;
    const uint32_t mp = c->prevCount, total = m + mp;
    const int nb = bodyBits(c->n);
    c->prevStage.reserve(m ? m : 1);
    if (m) kScatterBack<<<blocks(m), kThreads, 0, s>>>(c->solve.p, c->order.p, m, c->prevStage.p);
    c->carry.reserve(mp + 1); c->carryRank.reserve(mp + 1);
    c->cacheKeys.reserve(total + 1); c->cacheVals.reserve(total + 1);
    if (mp) {
        kCarryFlags<<<blocks(mp), kThreads, 0, s>>>(c->prev.p, mp, nb, c->evAppKeys.p, k + cEvInert, c->carry.p);
        HIPCK(vxhost::exclusiveScan(c->carry.p, c->carryRank.p, mp, k + cCarried, tempFor(c, vxhost::scanTempBytes(mp)), s));
    } else {
        kFill32<<<1, 64, 0, s>>>(k + cCarried, 1, 0u);
    }
    // Ordered as std::sort orders them; the carried count is read first (a
    // sleeping world is small, so the round trip is cheap).
    if (total) {
        kCacheSolveKeys<<<blocks(total), kThreads, 0, s>>>(c->solve.p, c->order.p, m, c->prev.p, mp, nb, c->carry.p,
                                                           c->carryRank.p, c->cacheKeys.p, c->cacheVals.p);
        HIPCK(hipMemcpyAsync(c->hostCounters + cCarried, k + cCarried, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        exactOrder(c, c->cacheKeys.p, c->cacheVals.p, m + c->hostCounters[cCarried]);
    }
    c->prevNext.reserve(total ? total : 1);
    if (total)
        kCacheGather<<<blocks(total), kThreads, 0, s>>>(c->prevStage.p, m, c->prev.p, mp, c->cacheVals.p, k + cCarried,
                                                        c->prevNext.p);
    HIPCK(hipGetLastError());
}

void finishCache(vxhip_ctx* c, uint32_t m) {
    std::swap(c->prev.p, c->prevNext.p);
    std::swap(c->prev.cap, c->prevNext.cap);
    c->prevCount = m + c->hostCounters[cCarried];
    c->prevExact = true;
}

void finishSleep(vxhip_ctx* c) {
    const uint32_t* h = c->hostCounters;
    vx_sleep_stats& st = c->sleepStats;
    st.totalDynamicBodies = h[cSlDynamic];
    st.awakeBodies = h[cSlDynamic + 1];
    st.drowsyBodies = h[cSlDynamic + 2];
    st.sleepingBodies = h[cSlDynamic + 3];
    st.totalSleepTransitions += h[cSlDynamic + 4];
    st.totalDrowsyTransitions += h[cSlDynamic + 5];
    st.totalWakeTransitions += h[cSlWoke];
    st.islandCount = h[cSlIslands];
    st.sleepingIslandCount = h[cSlIslands + 1];
    st.drowsyIslandCount = h[cSlIslands + 2];
    st.awakeIslandCount = h[cSlIslands + 3];
    c->anyAsleep = st.sleepingBodies > 0;
}

// One full device step of the resident world.
// Captures one later substep (integration, the velocity solve without the
// impulse store, the joint passes, transforms) on the context's stream and turns it into the
// context's executable graph: updated in place when the new capture has the
// old one's shape, instantiated anew otherwise.  The capture only records:
// every buffer it names was sized by the step's first substep, and no fill in
// it is a memset node.
void captureSubstep(vxhip_ctx* c, uint32_t K, V3 g, float h) {
    hipStream_t s = c->stream;
    const uint32_t n = c->n;
    uint32_t before[VXHIP_LAUNCH_KINDS];
    for (int l = 0; l < VXHIP_LAUNCH_KINDS; ++l) before[l] = c->stats.launches[l];
    HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t graph = nullptr;
    try {
        kIntegrate<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, g, h);
        substepSolve(c, c->solve.p, K, h, c->cfg.options, false, false);
        substepJoints(c, h, false);
        kTransforms<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, h);
    } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    HIPCK(hipStreamEndCapture(s, &graph));
    for (int l = 0; l < VXHIP_LAUNCH_KINDS; ++l) {
        c->graphLaunches[l] = c->stats.launches[l] - before[l];
        c->stats.launches[l] = before[l];
    }
    if (c->subExec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult why;
        if (hipGraphExecUpdate(c->subExec, graph, &bad, &why) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(c->subExec);
            c->subExec = nullptr;
        }
    }
    if (!c->subExec) {
        const hipError_t e = hipGraphInstantiate(&c->subExec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCK(e);
    } else {
        HIPCK(hipGraphDestroy(graph));
    }
    ++c->graphCaptures;
}

void worldStep(vxhip_ctx* c, float dt) {
    injectFailureForTesting();
    const vxhip_config& cfg = c->cfg;
    const int S = cfg.substeps > 0 ? cfg.substeps : 1;
    const float h = dt / S;
    const V3 g{cfg.gravity[0], cfg.gravity[1], cfg.gravity[2]};
    const uint32_t n = c->n;
    hipStream_t s = c->stream;
    c->prevPos.reserve(n); c->prevQ.reserve(n);
    uint32_t* k = c->counters.p;
    for (uint32_t& l : c->stats.launches) l = 0;
    markStage(c, sBegin);
    kResetCounters<<<1, 128, 0, s>>>(k, false);
    DiffStepRec* cap = c->diffCap ? c->diffCap : (c->recordOn ? &c->record : nullptr);
    if (cap) {
        cap->start.reserve(n);
        HIPCK(hipMemcpyAsync(cap->start.p, c->bodies.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToDevice, s));
        cap->prevCount = c->prevCount;
    }
    kBeginStep<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, g, h, c->prevPos.p, c->prevQ.p, 1);
    HIPCK(hipGetLastError());
    if (c->anyHigh && c->toiSettings.enabled && c->toiSettings.splitting && n) {  // world.cpp:4015-4025
        c->toiStarts.reserve(n);
        HIPCK(hipMemcpyAsync(c->toiStarts.p, c->bodies.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToDevice, s));
    }
    markStage(c, sBroad);
    const uint32_t np = broadPhase(c, dt);
    markStage(c, sNarrow);
    const uint32_t m = jointContactsOff(c, np, narrowPhase(c, np, dt));
    markStage(c, sMatch);
    if (c->anyAsleep && m) kWakePass<<<1, kWakeThreads, 0, s>>>(c->bodies.p, c->contacts.p, m);
    // Ties of the match are counted here and read with the colouring's
    // counters (colours do not depend on impulses); the match itself runs
    // after, against a cache in the reference's exact order when a tie needs it.
    const bool tieCheck = m && c->prevCount && !c->prevExact;
    if (tieCheck)
        kWarmTies<<<blocks(m), kThreads, 0, s>>>(c->contacts.p, m, c->prev.p, c->prevCount, k + cWarmTies, cap != nullptr);
    markStage(c, sColor);
    const uint32_t K = colorContacts(c, c->contacts.p, m);
    if (cap) {
        cap->matchCanon.reserve(m ? m : 1);
        fill32(c, cap->matchCanon.p, m, 0xffffffffu);
    }
    if (m && c->prevCount) {
        if (tieCheck && c->hostCounters[cWarmTies]) exactPrev(c);
        kWarmMatch<<<blocks(m), kThreads, 0, s>>>(c->contacts.p, m, c->prev.p, c->prevCount,
                                                  cap ? c->prevRank.p : nullptr, cap ? cap->matchCanon.p : nullptr);
    }
    colorOrder(c, m, K);
    c->solve.reserve(m ? m : 1);
    if (m) kGather<<<blocks(m), kThreads, 0, s>>>(c->contacts.p, c->order.p, m, c->solve.p);
    if (cap && m) {  // the step as the adjoint replays it
        cap->cs.reserve(m); cap->match.reserve(m); cap->colorOf.reserve(m);
        HIPCK(hipMemcpyAsync(cap->cs.p, c->solve.p, (size_t)m * sizeof(vx_contact), hipMemcpyDeviceToDevice, s));
        kGatherU32<<<blocks(m), kThreads, 0, s>>>(cap->matchCanon.p, c->order.p, m, cap->match.p);
        kColorOfSolve<<<blocks(m), kThreads, 0, s>>>(c->color.p, c->order.p, m, cap->colorOf.p);
    }
    c->solveCount = m;
    markStage(c, sSolve);
    if (c->jointCount && c->anyAsleep)  // backend_cuda.cu:704-708 (nothing to wake when no body sleeps)
        kJointWake<<<1, 1, 0, s>>>(c->bodies.p, c->joints.p, c->jointCount);
    // Substeps 2..S launch the same kernels with the same arguments once the
    // first has set up the step's islands and rows: with graphSubsteps they are
    // captured once (as substep 2) and replayed from one executable graph,
    // updated in place from step to step while its shape holds.
    const bool graph = c->graphSubsteps && m && S > 1 && !c->profile && !cap && !c->bigCount;
    for (int st = 0; st < S; ++st) {
        if (graph && st > 0) {
            if (st == 1) captureSubstep(c, K, g, h);
            else ++c->graphReplays;
            HIPCK(hipGraphLaunch(c->subExec, s));
            for (int l = 0; l < VXHIP_LAUNCH_KINDS; ++l) c->stats.launches[l] += c->graphLaunches[l];
            if (st == S - 1)
                kStoreImpulses<<<blocks(m), kThreads, 0, s>>>(c->solve.p, m, c->sImp.p, c->solvePos.p);
            continue;
        }
        if (st > 0) kIntegrate<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, g, h);
        if (m) substepSolve(c, c->solve.p, K, h, cfg.options, st == 0, st == S - 1);
        substepJoints(c, h, st == 0);
        kTransforms<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, h);
    }
    HIPCK(hipGetLastError());
    markStage(c, sCcd);
    kResetCounters<<<1, 128, 0, s>>>(k);
    if (c->jointCount) {  // World::finishBrokenJoints (world.cpp:3765-3783)
        c->jointBreaks.reserve(c->jointCount);
        kJointBroken<<<blocks(c->jointCount), kThreads, 0, s>>>(c->joints.p, c->jointCount, h, k + cJointsBroken);
        kJointFinish<<<1, 1, 0, s>>>(c->joints.p, c->jointCount, h, c->jointBreaks.p, k + cJointsLeft, k + cJointBreaks,
                                     k + cJointsBroken);
    }
    highToi(c, m, dt);
    ccdNet(c, np, dt);
    markStage(c, sPosition);
    if (m) positionCorrection(c, c->solve.p, K, h, cfg.options);
    markStage(c, sFinish);
    // With sleeping on, or a body asleep, the events' pair history also
    // decides which inert pairs keep their cached contacts.
    const bool carry = c->sleepOn || c->anyAsleep;
    const bool events = c->eventsEnabled || carry;
    if (events) contactEvents(c, m, dt);
    if (c->sleepOn) sleepUpdate(c, m, dt);
    kClearForces<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n);
    if (carry) {
        buildCache(c, m);
    } else {
        c->prev.reserve(m ? m : 1);
        c->prevRank.reserve(m ? m : 1);
        if (m) kScatterBack<<<blocks(m), kThreads, 0, s>>>(c->solve.p, c->order.p, m, c->prev.p);
        if (m) kRankBack<<<blocks(m), kThreads, 0, s>>>(c->order.p, m, c->prevRank.p);
    }
    HIPCK(hipGetLastError());
    c->solveCount = m;
    c->numColors = K;
    c->lastIterations = (uint32_t)(2 * cfg.options.velocityIterations * S);
    markStage(c, sEnd);
    readCounters(c);
    if (c->hostCounters[cColorError]) fail(VXHIP_ERR_DEVICE_LOST, "vxhip: CCD dataflow did not converge");
    if (events) finishEvents(c);
    if (!c->eventsEnabled) c->eventCount = 0;
    if (carry) {
        finishCache(c, m);
    } else {
        c->prevCount = m;
        c->prevExact = false;
    }
    c->stats.warmTies = c->hostCounters[cWarmTies];
    if (c->sleepOn) finishSleep(c);
    c->jointBreakCount = 0;
    if (c->jointCount) {
        c->jointBreakCount = c->hostCounters[cJointBreaks];
        const uint32_t left = c->hostCounters[cJointsLeft];
        if (left != c->jointCount) {  // joints broke: the host list follows, colours are redone
            c->hostJoints.resize(left);
            if (left)
                HIPCK(hipMemcpyAsync(c->hostJoints.data(), c->joints.p, (size_t)left * sizeof(vx_joint),
                                     hipMemcpyDeviceToHost, c->stream));
            HIPCK(hipStreamSynchronize(c->stream));
            prepareJoints(c);
        }
    }
    if (c->hostCounters[cBatchError]) fail(VXHIP_ERR_DEVICE_LOST, "vxhip: island batch exceeds its LDS capacity");
    collectProfile(c);
    c->stats.bodies = n;
    c->stats.pairs = np;
    c->stats.contacts = m;
    c->stats.colors = K;
    c->stats.ccdCandidates = c->hostCounters[cCcdExamined];
    c->stats.ccdRescued = c->hostCounters[cCcdRescued];
    c->stats.ccdFlagged = c->hostCounters[cCcdFlagged];
    c->stats.ccdRelevant = c->hostCounters[cCcdRelevant];
    c->stats.velocityIterations = c->lastIterations;
    c->stats.islands = m ? c->islandCount : 0;
    c->stats.islandContacts = m ? c->smallCount : 0;
    c->stats.bigIslandContacts = m ? c->bigCount : 0;
    c->stats.largestIsland = m ? c->largestIsland : 0;
    c->stats.maxBatchPasses = m ? c->maxBatchPasses : 0;
    c->stats.multiToiEvents = c->toiEventCount;
    c->stats.graphCaptures = c->graphCaptures;
    c->stats.graphReplays = c->graphReplays;
    if (cap) {
        cap->m = m;
        cap->ccdRescued = c->hostCounters[cCcdRescued];
    }
}

// -------------------------------------------------- differentiable form --
// Rollouts of the resident world whose gradients the adjoint kernels of
// vx_diff.hpp compute.  The forward runs the ordinary device step, keeping a
// checkpoint of everything a step reads from the one before (bodies and the
// warm-start cache) every `diffEvery` steps.  The backward goes over the
// checkpoint segments from the last: it replays a segment with the ordinary
// step, recording each step's contacts (solve order, warm-start impulses),
// colours and warm-start matches; then, step by step from the segment's end,
// it replays the step's substeps with the per-world tape kernels (the same
// rows, so bit for bit the recorded trajectory) and runs the adjoint kernels
// over them in reverse.
constexpr uint32_t kDiffMaxSubsteps = 8;

__global__ void kFillF4(float4* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
__global__ void kFillF(float* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}
// Bodies whose replayed state (pose and velocities) differs from the
// recorded one, bitwise.
__global__ void kDiffCompare(const vx_body* __restrict__ a, const vx_body* __restrict__ b, uint32_t n,
                             uint32_t* __restrict__ count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* x = reinterpret_cast<const uint32_t*>(&a[i]);
    const uint32_t* y = reinterpret_cast<const uint32_t*>(&b[i]);
    bool same = true;
    for (int w = 0; w < 13; ++w) same &= x[w] == y[w];  // position, orientation, velocity, angular velocity
    if (!same) atomicAdd(count, 1u);
}

void zeroF4(vxhip_ctx* c, Buf<float4>& b, uint32_t n) {
    b.reserve(n ? n : 1);
    if (n) kFillF4<<<blocks(n), kThreads, 0, c->stream>>>(b.p, n);
}

// What the differentiable form admits: dynamic bodies of isotropic inverse
// inertia in their principal frame, geometric-mean friction, no rolling or
// spinning resistance, no kinematic, locked, high-quality or sleeping bodies,
// no joints, the two-axis friction model; worlds in contiguous body ranges.
void diffValidate(vxhip_ctx* c) {
    if (!c->loaded) fail(VXHIP_ERR_INVALID, "vxhip: no world is loaded");
    if (c->sleepOn || c->jointCount || c->anyHigh)
        fail(VXHIP_ERR_UNAVAILABLE, "vxhip: differentiable rollouts support no sleeping, joints or multi-TOI bodies");
    if (c->cfg.options.frictionModel != VX_FRICTION_TWO_AXIS)
        fail(VXHIP_ERR_UNAVAILABLE, "vxhip: differentiable rollouts support the two-axis friction model only");
    if (c->cfg.substeps < 1 || (uint32_t)c->cfg.substeps > kDiffMaxSubsteps)
        fail(VXHIP_ERR_UNAVAILABLE, "vxhip: differentiable rollouts support 1 to 8 substeps");
    const uint32_t n = c->n;
    std::vector<vx_body> b(n);
    HIPCK(hipMemcpyAsync(b.data(), c->bodies.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToHost, c->stream));
    std::vector<uint32_t> worldOf;
    const bool batched = c->worldCount > 1 && c->worldN == n;
    if (batched) {
        worldOf.resize(n);
        HIPCK(hipMemcpyAsync(worldOf.data(), c->worldOf.p, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->stream));
    }
    HIPCK(hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < n; ++i) {
        const vx_body& x = b[i];
        if (x.frictionCombine != 1 || x.rollingFriction != 0.0f || x.spinningFriction != 0.0f)
            fail(VXHIP_ERR_UNAVAILABLE, "vxhip: differentiable rollouts need geometric-mean friction without rolling "
                                        "or spinning resistance");
        if (x.motionType == 0) continue;  // static
        const bool iso = x.invInertia.x == x.invInertia.y && x.invInertia.y == x.invInertia.z &&
                         x.inertiaOrientation.x == 0.0f && x.inertiaOrientation.y == 0.0f &&
                         x.inertiaOrientation.z == 0.0f && x.inertiaOrientation.w == 1.0f;
        if (x.motionType != 2 || !iso || x.fixedRotation || x.ccdTuning.quality == 3 || x.ccdTuning.quality == 2 ||
            x.invMass <= 0.0f || x.asleep)
            fail(VXHIP_ERR_UNAVAILABLE, "vxhip: differentiable rollouts need dynamic bodies of isotropic inertia "
                                        "(spheres, cubes) of Medium or Low motion quality");
    }
    c->diffW = batched ? c->worldCount : 1u;
    c->diffBodyStartHost.assign(c->diffW + 1, 0u);
    if (batched) {
        for (uint32_t i = 0; i < n; ++i) {
            if (worldOf[i] >= c->diffW || (i && worldOf[i] < worldOf[i - 1]))
                fail(VXHIP_ERR_UNAVAILABLE, "vxhip: differentiable rollouts need each world's bodies contiguous");
            c->diffBodyStartHost[worldOf[i] + 1] = i + 1;
        }
        for (uint32_t w = 1; w <= c->diffW; ++w)
            c->diffBodyStartHost[w] = std::max(c->diffBodyStartHost[w], c->diffBodyStartHost[w - 1]);
    } else {
        c->diffBodyStartHost[1] = n;
    }
    c->diffBodyStart.reserve(c->diffW + 1);
    HIPCK(hipMemcpyAsync(c->diffBodyStart.p, c->diffBodyStartHost.data(), (c->diffW + 1) * sizeof(uint32_t),
                         hipMemcpyHostToDevice, c->stream));
}

void diffSave(vxhip_ctx* c, DiffCheckpoint& cp) {
    hipStream_t s = c->stream;
    cp.bodies.reserve(c->n);
    HIPCK(hipMemcpyAsync(cp.bodies.p, c->bodies.p, (size_t)c->n * sizeof(vx_body), hipMemcpyDeviceToDevice, s));
    cp.prevCount = c->prevCount;
    cp.prevExact = c->prevExact;
    cp.prev.reserve(c->prevCount ? c->prevCount : 1);
    cp.prevRank.reserve(c->prevCount ? c->prevCount : 1);
    if (c->prevCount) {
        HIPCK(hipMemcpyAsync(cp.prev.p, c->prev.p, (size_t)c->prevCount * sizeof(vx_contact), hipMemcpyDeviceToDevice, s));
        HIPCK(hipMemcpyAsync(cp.prevRank.p, c->prevRank.p, (size_t)c->prevCount * sizeof(uint32_t),
                             hipMemcpyDeviceToDevice, s));
    }
}

void diffRestore(vxhip_ctx* c, const DiffCheckpoint& cp) {
    hipStream_t s = c->stream;
    HIPCK(hipMemcpyAsync(c->bodies.p, cp.bodies.p, (size_t)c->n * sizeof(vx_body), hipMemcpyDeviceToDevice, s));
    c->prev.reserve(cp.prevCount ? cp.prevCount : 1);
    c->prevRank.reserve(cp.prevCount ? cp.prevCount : 1);
    if (cp.prevCount) {
        HIPCK(hipMemcpyAsync(c->prev.p, cp.prev.p, (size_t)cp.prevCount * sizeof(vx_contact), hipMemcpyDeviceToDevice, s));
        HIPCK(hipMemcpyAsync(c->prevRank.p, cp.prevRank.p, (size_t)cp.prevCount * sizeof(uint32_t),
                             hipMemcpyDeviceToDevice, s));
    }
    c->prevCount = cp.prevCount;
    c->prevExact = cp.prevExact;
}

// The solve positions of one recorded step grouped by world (kDiffSweeps'
// `list` and `wStart`).
void diffGroup(vxhip_ctx* c, const DiffStepRec& r) {
    hipStream_t s = c->stream;
    const uint32_t m = r.m, W = c->diffW;
    c->dList.reserve(m ? m : 1); c->dListTmp.reserve(m ? m : 1); c->dWStart.reserve(W + 1); c->dWCount.reserve(W);
    if (!m) { fill32(c, c->dWStart.p, W + 1, 0u); return; }
    if (W == 1) {
        kIota<<<blocks(m), kThreads, 0, s>>>(c->dList.p, m);
        const uint32_t ws[2] = {0u, m};
        HIPCK(hipMemcpyAsync(c->dWStart.p, ws, sizeof ws, hipMemcpyHostToDevice, s));
        HIPCK(hipStreamSynchronize(s));
        return;
    }
    c->dWKeys.reserve(m); c->dWKeysTmp.reserve(m);
    fill32(c, c->dWCount.p, W, 0u);
    kDiffWorldKeys<<<blocks(m), kThreads, 0, s>>>(r.cs.p, m, c->worldOf.p, c->dWKeys.p, c->dList.p, c->dWCount.p);
    HIPCK(vxhost::radixSort(c->dWKeys.p, c->dList.p, c->dWKeysTmp.p, c->dListTmp.p, m, vxhost::bitsFor(W - 1),
                            tempFor(c, vxhost::radixTempBytes(m)), s));
    HIPCK(vxhost::exclusiveScan(c->dWCount.p, c->dWStart.p, W, c->dWStart.p + W, tempFor(c, vxhost::scanTempBytes(W)), s));
}

// Reverse of one recorded step: on entry the adjoints hold those of the
// step's end state and diffLamBar those of its contacts' final impulses; on
// return, those of its start state, and diffLamPrev those of the previous
// step's final impulses.
void diffBackStep(vxhip_ctx* c, const DiffStepRec& r, const vx_body* next) {
    hipStream_t s = c->stream;
    const uint32_t n = c->n, m = r.m, W = c->diffW;
    const int S = c->cfg.substeps;
    const float h = c->diffDt / (float)S;
    const V3 g{c->cfg.gravity[0], c->cfg.gravity[1], c->cfg.gravity[2]};
    const vx_solver_options& opt = c->cfg.options;
    const int sweeps = 2 * opt.velocityIterations;
    uint32_t* over = c->diffOverflow.p;
    diffGroup(c, r);
    // Forward replay with the tape.
    c->diffWork.reserve(n); c->diffPost.reserve(n);
    c->dVel.reserve(n); c->dAng.reserve(n); c->dInvI.reserve(n); c->dFrame.reserve(n);
    c->dMeta.reserve(m ? m : 1); c->dImp.reserve(2 * (size_t)(m ? m : 1)); c->dSide.reserve(8 * (size_t)(m ? m : 1));
    c->dIota.reserve(m ? m : 1);
    c->diffCTape.reserve((size_t)(m ? m : 1) * (opt.positionIterations > 0 ? opt.positionIterations : 1));
    c->contactsTmp.reserve(m ? m : 1);
    if (m) {
        kIota<<<blocks(m), kThreads, 0, s>>>(c->dIota.p, m);
        HIPCK(hipMemcpyAsync(c->contactsTmp.p, r.cs.p, (size_t)m * sizeof(vx_contact), hipMemcpyDeviceToDevice, s));
    }
    HIPCK(hipMemcpyAsync(c->diffWork.p, r.start.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToDevice, s));
    kBeginStep<<<blocks(n), kThreads, 0, s>>>(c->diffWork.p, n, g, h, nullptr, nullptr, 1);
    SolveArrays sa{c->dVel.p, c->dAng.p, c->dInvI.p, c->dFrame.p, nullptr, c->dMeta.p, c->dImp.p, c->dSide.p, c->dIota.p};
    for (int st = 0; st < S; ++st) {
        if (st > 0) kIntegrate<<<blocks(n), kThreads, 0, s>>>(c->diffWork.p, n, g, h);
        c->diffSub[st].reserve(n); c->diffPre[st].reserve(n);
        c->diffRows[st].reserve(8 * (size_t)(m ? m : 1));
        c->diffTape[st].reserve(4 * (size_t)(sweeps + 1) * (m ? m : 1));
        HIPCK(hipMemcpyAsync(c->diffSub[st].p, c->diffWork.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToDevice, s));
        kGatherSolveBodies<<<blocks(n), kThreads, 0, s>>>(c->diffWork.p, n, sa);
        sa.rows = c->diffRows[st].p;
        if (m) {
            kPrepareRows<<<blocks(m), kThreads, 0, s>>>(c->diffWork.p, c->contactsTmp.p, m, h, sa);
            if (st == 0) kLoadImpulses<<<blocks(m), kThreads, 0, s>>>(r.cs.p, m, sa.imp, sa.pos);
            kDiffSweeps<<<W, kDiffThreads, 0, s>>>(sa, c->dList.p, c->dWStart.p, r.colorOf.p, m, sweeps, st == 0,
                                                   c->diffTape[st].p, over);
            kScatterVelocities<<<blocks(n), kThreads, 0, s>>>(c->diffWork.p, n, sa);
        }
        HIPCK(hipMemcpyAsync(c->diffPre[st].p, c->diffWork.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToDevice, s));
        kTransforms<<<blocks(n), kThreads, 0, s>>>(c->diffWork.p, n, h);
    }
    HIPCK(hipMemcpyAsync(c->diffPost.p, c->diffWork.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToDevice, s));
    if (m && opt.positionIterations > 0)
        kDiffCorrection<<<W, kDiffThreads, 0, s>>>(c->diffWork.p, c->contactsTmp.p, c->dList.p, c->dWStart.p,
                                                   r.colorOf.p, m, opt.positionIterations, h, opt, c->diffCTape.p, over);
    kDiffCompare<<<blocks(n), kThreads, 0, s>>>(c->diffWork.p, next, n, over + 1);
    HIPCK(hipGetLastError());
    // Reverse.
    float4 *aX = c->diffAdjX.p, *aQ = c->diffAdjQ.p, *aV = c->diffAdjV.p, *aW = c->diffAdjW.p, *aP = c->diffAdjParam.p;
    c->diffRowBar.reserve(4 * (size_t)(m ? m : 1));
    if (m) zeroF4(c, c->diffRowBar, 4 * m);
    if (m && opt.positionIterations > 0)
        kDiffCorrectionAdj<<<W, kDiffThreads, 0, s>>>(c->diffPost.p, c->contactsTmp.p, c->dList.p, c->dWStart.p,
                                                      r.colorOf.p, m, opt.positionIterations, c->diffCTape.p, aX, aQ,
                                                      aP, over);
    for (int st = S - 1; st >= 0; --st) {
        kDiffTransformAdj<<<blocks(n), kThreads, 0, s>>>(c->diffPre[st].p, n, h, aX, aQ, aV, aW);
        if (!m) continue;
        kGatherSolveBodies<<<blocks(n), kThreads, 0, s>>>(c->diffSub[st].p, n, sa);
        sa.rows = c->diffRows[st].p;
        kDiffSweepsAdj<<<W, kDiffThreads, 0, s>>>(sa, c->dList.p, c->dWStart.p, r.colorOf.p, m, sweeps, st == 0, h,
                                                  c->diffTape[st].p, aV, aW, c->diffLamBar.p, c->diffRowBar.p, over);
        kDiffPrefixAdj<<<blocks(m), kThreads, 0, s>>>(sa, r.cs.p, m, c->diffSub[st].p, h, c->diffRowBar.p, aX, aQ, aP);
    }
    kDiffNormalizeAdj<<<blocks(n), kThreads, 0, s>>>(r.start.p, n, aQ);
    // The warm-start impulses were the matched contacts' final impulses.
    zeroF4(c, c->diffLamPrev, r.prevCount);
    if (m && r.prevCount)
        kDiffLambdaCarry<<<blocks(m), kThreads, 0, s>>>(r.match.p, m, c->diffLamBar.p, c->diffLamPrev.p);
    HIPCK(hipGetLastError());
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return VXHIP_OK;
    } catch (const Failure& e) {
        return e.code;
    } catch (const std::exception& e) {
        gLastError = e.what();
        return VXHIP_ERR_INVALID;
    }
}

vxhip_ctx* newContext() {
    vxhip_ctx* c = new vxhip_ctx();
    HIPCK(hipGetDevice(&c->device));
    HIPCK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
        c->cuCount = (uint32_t)prop.multiProcessorCount;
    c->batchQueue.reserve(2);
    kFill32<<<1, 64, 0, c->stream>>>(c->batchQueue.p, 2, 0u);
    HIPCK(hipGetLastError());
    c->counters.reserve(cCount);
    c->extHist.reserve(kExtBuckets);
    kFill32<<<1, kExtBuckets, 0, c->stream>>>(c->extHist.p, kExtBuckets, 0u);
    HIPCK(hipGetLastError());
    HIPCK(hipHostMalloc((void**)&c->hostCounters, cCount * sizeof(uint32_t), hipHostMallocDefault));
    return c;
}

}  // namespace

// ------------------------------------------------ allocation-free sort --
// vxhip_std_sort_noalloc works in a fixed device array instead of memory it
// allocates: 256 bytes for its flag word, then the sort's working memory.
// Every device holds its own copy at its own address, found once per device
// with hipGetSymbolAddress; the array's name is never used as an address in
// host code.  Calls on one device share the array (and keep the last call's
// contents): calls that overlap on one device race.
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
constexpr uint32_t kNoAllocMaxKeys = 262144, kNoAllocMaxSegments = 4096, kNoAllocFlagSlot = cCount - 1;
static_assert(kNoAllocFlagSlot > cWarmTies, "the flag word's host slot is free");

}  // namespace

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
void* gScratch[kMaxScratchDevices] = {};
std::mutex gScratchMutex;

// The device address of `scratch` on device `dev` (current), once per device.
uint8_t* scratchOn(int dev) {
    if (dev < 0 || dev >= kMaxScratchDevices) fail(VXHIP_ERR_UNAVAILABLE, "vxhip: device index out of range");
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) fail(classify(e), std::string("vxhip: scratch array address: ") + hipGetErrorString(e));
        gScratch[dev] = p;
    }
    return (uint8_t*)gScratch[dev];
}

}  // namespace

// ------------------------------------------------------------------ C ABI --
extern "C" {

const char* vxhip_last_error(void) { return gLastError.c_str(); }

int vxhip_device_count(int* count) {
    return guarded([&] {
        if (!count) fail(VXHIP_ERR_INVALID, "vxhip: null count");
        *count = 0;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) return;
        for (int d = 0; d < n; ++d) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, d) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
                ++*count;
        }
    });
}

int vxhip_create(vxhip_ctx** out) {
    return guarded([&] {
        if (!out) fail(VXHIP_ERR_INVALID, "vxhip: null output");
        *out = nullptr;
        int count = 0;
        vxhip_device_count(&count);
        if (count == 0) fail(VXHIP_ERR_UNAVAILABLE, "vxhip: no gfx950 device");
        *out = newContext();
    });
}

void vxhip_destroy(vxhip_ctx* c) {
    if (!c) return;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) std::fprintf(stderr, "vxhip cleanup: %s\n", hipGetErrorString(e));
    // Device buffers are released by their destructors below.
    for (hipEvent_t e : c->stageEv) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->launchEv) (void)hipEventDestroy(e);
    if (c->subExec) (void)hipGraphExecDestroy(c->subExec);
    if (c->forkEv) (void)hipEventDestroy(c->forkEv);
    if (c->joinEv) (void)hipEventDestroy(c->joinEv);
    if (c->hostCounters) (void)hipHostFree(c->hostCounters);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int vxhip_world_load(vxhip_ctx* c, const vx_body* bodies, uint32_t n, const vx_soup* soup, const vxhip_config* cfg) {
    return guarded([&] {
        if (!c || (!bodies && n) || !cfg) fail(VXHIP_ERR_INVALID, "vxhip: invalid world arguments");
        validateScene(bodies, n, soup);
        if (c->sleepOn && n > kSleepMaxBodies)
            fail(VXHIP_ERR_CAPACITY, "vxhip: device sleeping is limited to 256 bodies in this build");
        c->cfg = *cfg;
        uploadSoup(c, soup);
        setBodies(c, bodies, n);
        c->prevCount = 0;
        c->prevExact = false;
        c->solveCount = 0;
        c->worldCount = 1;
        c->worldN = 0;
        c->eventCount = 0;
        c->evPrevCount = 0;
        c->hostJoints.clear();
        c->jointCount = c->jointColors = c->jointExcludedCount = c->jointBreakCount = 0;
        HIPCK(hipStreamSynchronize(c->stream));
        c->loaded = true;
    });
}

int vxhip_world_step(vxhip_ctx* c, float dt, uint32_t steps) {
    return guarded([&] {
        if (!c || !c->loaded) fail(VXHIP_ERR_INVALID, "vxhip: no world loaded");
        if (!(dt >= 0.0f) || dt > 1e30f) fail(VXHIP_ERR_INVALID, "vxhip: timestep must be finite and non-negative");
        if (c->sleepOn && c->n > kSleepMaxBodies)  // bodies replaced since sleeping was switched on
            fail(VXHIP_ERR_CAPACITY, "vxhip: device sleeping is limited to 256 bodies in this build");
        if (dt == 0.0f) return;
        for (uint32_t i = 0; i < steps; ++i) worldStep(c, dt);
    });
}

int vxhip_world_set_worlds(vxhip_ctx* c, const uint32_t* worldOf, uint32_t worldCount) {
    return guarded([&] {
        if (!c || !c->loaded) fail(VXHIP_ERR_INVALID, "vxhip: no world loaded");
        if (worldCount == 0 || (worldCount > 1 && !worldOf)) fail(VXHIP_ERR_INVALID, "vxhip: invalid world ids");
        if (worldCount > 1) {
            for (uint32_t i = 0; i < c->n; ++i)
                if (worldOf[i] >= worldCount) fail(VXHIP_ERR_INVALID, "vxhip: world id out of range");
            c->worldOf.reserve(c->n);
            HIPCK(hipMemcpyAsync(c->worldOf.p, worldOf, (size_t)c->n * sizeof(uint32_t), hipMemcpyHostToDevice,
                                 c->stream));
            HIPCK(hipStreamSynchronize(c->stream));
        }
        c->worldCount = worldCount;
        c->worldN = c->n;
        c->worldsContiguous = worldCount > 1;
        for (uint32_t i = 1; worldCount > 1 && i < c->n; ++i)
            if (worldOf[i] < worldOf[i - 1]) c->worldsContiguous = false;
    });
}

int vxhip_world_set_island_limit(vxhip_ctx* c, uint32_t limit) {
    return guarded([&] {
        if (!c || limit == 0) fail(VXHIP_ERR_INVALID, "vxhip: invalid island limit");
        c->islandLimit = limit;
    });
}

int vxhip_world_set_tuning(vxhip_ctx* c, const vxhip_tuning* t) {
    return guarded([&] {
        if (!c || !t) fail(VXHIP_ERR_INVALID, "vxhip: invalid tuning arguments");
        if (t->batchThreads && t->batchThreads != 256 && t->batchThreads != 512)
            fail(VXHIP_ERR_INVALID, "vxhip: batchThreads must be 256 or 512");
        if (t->persistentBatches > 2 || t->longestFirst > 2 || t->sideStream > 2 || t->graphSubsteps > 2)
            fail(VXHIP_ERR_INVALID, "vxhip: switch fields take 0 (default), 1 (on) or 2 (off)");
        c->islandLimit = t->islandLimit ? t->islandLimit : 4096u;
        c->batchThreads = t->batchThreads ? t->batchThreads : 256u;
        c->persistentBatches = t->persistentBatches != 2;
        c->longestFirst = t->longestFirst != 2;
        c->colorLaunchRows = t->colorLaunchRows ? t->colorLaunchRows : 1024u;
        c->sideOn = t->sideStream != 2;
        c->graphSubsteps = t->graphSubsteps == 1;
        const uint32_t packMax = t->jointPackMax ? t->jointPackMax : kJointPackMax;
        if (packMax != c->jointPackMax) {
            c->jointPackMax = packMax;
            if (c->jointCount) prepareJoints(c);  // the packs follow the limit
        }
    });
}

int vxhip_world_get_tuning(vxhip_ctx* c, vxhip_tuning* out) {
    return guarded([&] {
        if (!c || !out) fail(VXHIP_ERR_INVALID, "vxhip: invalid tuning arguments");
        *out = vxhip_tuning{};
        out->islandLimit = c->islandLimit;
        out->batchThreads = c->batchThreads;
        out->persistentBatches = c->persistentBatches ? 1u : 2u;
        out->longestFirst = c->longestFirst ? 1u : 2u;
        out->colorLaunchRows = c->colorLaunchRows;
        out->sideStream = c->sideOn ? 1u : 2u;
        out->graphSubsteps = c->graphSubsteps ? 1u : 2u;
        out->jointPackMax = c->jointPackMax;
    });
}

int vxhip_world_sync(vxhip_ctx* c) {
    return guarded([&] { HIPCK(hipStreamSynchronize(c->stream)); });
}

int vxhip_world_download_bodies(vxhip_ctx* c, vx_body* out, uint32_t n) {
    return guarded([&] {
        if (!c || n != c->n) fail(VXHIP_ERR_INVALID, "vxhip: body count mismatch");
        if (n) HIPCK(hipMemcpyAsync(out, c->bodies.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int vxhip_world_upload_bodies(vxhip_ctx* c, const vx_body* in, uint32_t n) {
    return guarded([&] {
        if (!c || n != c->n) fail(VXHIP_ERR_INVALID, "vxhip: body count mismatch");
        setBodies(c, in, n);
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int vxhip_world_download_contacts(vxhip_ctx* c, vx_contact* out, uint32_t cap, uint32_t* count) {
    return guarded([&] {
        if (!c || !count) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        *count = c->solveCount;
        if (c->solveCount > cap) fail(VXHIP_ERR_CAPACITY, "vxhip: contact buffer too small");
        if (c->solveCount)
            HIPCK(hipMemcpyAsync(out, c->solve.p, (size_t)c->solveCount * sizeof(vx_contact), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int vxhip_world_download_events(vxhip_ctx* c, vx_contact_event* out, uint32_t cap, uint32_t* count) {
    return guarded([&] {
        if (!c || !count || (!out && cap)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        *count = c->eventCount;
        if (c->eventCount > cap) fail(VXHIP_ERR_CAPACITY, "vxhip: event buffer too small");
        if (c->eventCount)
            HIPCK(hipMemcpyAsync(out, c->events.p, (size_t)c->eventCount * sizeof(vx_contact_event),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int vxhip_world_set_events(vxhip_ctx* c, int enable) {
    return guarded([&] {
        if (!c) fail(VXHIP_ERR_INVALID, "vxhip: null context");
        c->eventsEnabled = enable != 0;
        c->eventCount = 0;
        c->evPrevCount = 0;
    });
}

namespace {
void setJoints(vxhip_ctx* c, const vx_joint* joints, uint32_t count, bool anyCount) {
    if (!c || !c->loaded || (!joints && count)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
    if (!anyCount && count && count < kMinDeviceJoints)
        fail(VXHIP_ERR_UNAVAILABLE,
             "vxhip: fewer than 64 joints are solved by the host world (World::solveJoints), as with the CUDA backend");
    for (uint32_t k = 0; k < count; ++k)
        if (joints[k].type > VX_JOINT_GEAR || joints[k].a >= c->n || joints[k].b >= c->n)
            fail(VXHIP_ERR_INVALID, "vxhip: joint with an unknown type or body");
    c->hostJoints.assign(joints, joints + count);
    prepareJoints(c);
    c->jointBreakCount = 0;
}
}  // namespace

int vxhip_world_set_joints(vxhip_ctx* c, const vx_joint* joints, uint32_t count) {
    return guarded([&] { setJoints(c, joints, count, false); });
}

int vxhip_world_set_joints_all(vxhip_ctx* c, const vx_joint* joints, uint32_t count) {
    return guarded([&] { setJoints(c, joints, count, true); });
}

int vxhip_world_download_joints(vxhip_ctx* c, vx_joint* out, uint32_t cap, uint32_t* count) {
    return guarded([&] {
        if (!c || !count || (!out && cap)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        *count = c->jointCount;
        if (c->jointCount > cap) fail(VXHIP_ERR_CAPACITY, "vxhip: joint buffer too small");
        if (c->jointCount)
            HIPCK(hipMemcpyAsync(out, c->joints.p, (size_t)c->jointCount * sizeof(vx_joint), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int vxhip_world_download_joint_breaks(vxhip_ctx* c, vx_joint_break* out, uint32_t cap, uint32_t* count) {
    static_assert(sizeof(vx_joint_break) == sizeof(vxk::JointBreak), "joint break record");
    return guarded([&] {
        if (!c || !count || (!out && cap)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        *count = c->jointBreakCount;
        if (c->jointBreakCount > cap) fail(VXHIP_ERR_CAPACITY, "vxhip: break buffer too small");
        if (c->jointBreakCount)
            HIPCK(hipMemcpyAsync(out, c->jointBreaks.p, (size_t)c->jointBreakCount * sizeof(vx_joint_break),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int vxhip_world_set_multi_toi(vxhip_ctx* c, const vx_multi_toi_settings* v) {
    return guarded([&] {
        if (!c || !v) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        if (v->maxToiEventsPerBody == 0 || !std::isfinite(v->toiVelocityFloor) || v->toiVelocityFloor < 0.0f ||
            !std::isfinite(v->toiPenetrationBias) || v->toiPenetrationBias < 0.0f || v->maxTotalEventsPerStep == 0)
            fail(VXHIP_ERR_INVALID, "velox: multi-TOI settings are invalid");
        c->toiSettings = vxk::ToiSettings{v->enabled ? 1u : 0u, v->enableSubstepSplitting ? 1u : 0u,
                                          v->maxToiEventsPerBody, v->maxTotalEventsPerStep, v->toiVelocityFloor,
                                          v->toiPenetrationBias};
    });
}

int vxhip_world_query_multi_toi(vxhip_ctx* c, uint32_t body, float dt, vx_toi_hit* out, uint32_t cap,
                                uint32_t* count) {
    return guarded([&] {
        if (!c || !c->loaded || !count || (!out && cap)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        if (!std::isfinite(dt) || dt <= 0.0f)
            fail(VXHIP_ERR_INVALID, "velox: multi-TOI timestep must be finite and positive");
        if (body >= c->n) fail(VXHIP_ERR_INVALID, "vxhip: body index out of range");
        if (c->worldCount > 1) fail(VXHIP_ERR_UNAVAILABLE, "vxhip: queryMultiToi runs on a single-world context");
        const uint32_t n = c->n;
        hipStream_t s = c->stream;
        c->toiQueryBits.reserve(n);
        c->toiQueryHit.reserve(2 * (size_t)n);
        kToiQuery<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, c->soup, body, dt, c->toiSettings,
                                                c->toiQueryBits.p, c->toiQueryHit.p);
        HIPCK(hipGetLastError());
        std::vector<uint32_t> bits(n);
        std::vector<float4> hit(2 * (size_t)n);
        HIPCK(hipMemcpyAsync(bits.data(), c->toiQueryBits.p, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipMemcpyAsync(hit.data(), c->toiQueryHit.p, 2 * (size_t)n * sizeof(float4), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        // Sorted by (toi, body) as the reference sorts its hits (world.cpp:3174-3178);
        // non-negative times of impact order like their bit patterns.
        std::vector<uint64_t> order;
        for (uint32_t o = 0; o < n; ++o)
            if (bits[o] != vxk::kNone) order.push_back((uint64_t)bits[o] << 32 | o);
        std::sort(order.begin(), order.end());
        *count = (uint32_t)order.size();
        for (uint32_t k = 0; k < cap && k < order.size(); ++k) {
            const uint32_t o = (uint32_t)order[k];
            const float4 a = hit[2 * (size_t)o], b = hit[2 * (size_t)o + 1];
            out[k] = vx_toi_hit{b.w, o, {b.x, b.y, b.z}, {a.x, a.y, a.z}, a.w};
        }
    });
}

int vxhip_world_shift_origin(vxhip_ctx* c, const float offset[3]) {
    return guarded([&] {
        if (!c || !c->loaded || !offset) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        const V3 off{offset[0], offset[1], offset[2]};
        if (!std::isfinite(off.x) || !std::isfinite(off.y) || !std::isfinite(off.z))
            fail(VXHIP_ERR_INVALID, "velox: origin shift must be finite");
        hipStream_t s = c->stream;
        uint32_t* bad = c->counters.p + cShiftBad;
        for (int apply = 0; apply < 2; ++apply) {
            if (!apply) kFill32<<<1, 64, 0, s>>>(bad, 1, 0u);
            if (c->n) kShiftBodies<<<blocks(c->n), kThreads, 0, s>>>(c->bodies.p, c->n, off, apply, bad);
            if (c->vertexCount)
                kShiftPoints<<<blocks(c->vertexCount), kThreads, 0, s>>>(c->vertices.p, c->vertexCount, off, apply, bad);
            if (c->meshCount)
                kShiftBounds<vx_mesh><<<blocks(c->meshCount), kThreads, 0, s>>>(c->meshes.p, c->meshCount, off, apply, bad);
            if (c->nodeCount)
                kShiftBounds<vx_bvh_node><<<blocks(c->nodeCount), kThreads, 0, s>>>(c->nodes.p, c->nodeCount, off, apply,
                                                                                   bad);
            if (c->solveCount)
                kShiftContacts<<<blocks(c->solveCount), kThreads, 0, s>>>(c->solve.p, c->solveCount, off, apply, bad);
            if (c->prevCount)
                kShiftContacts<<<blocks(c->prevCount), kThreads, 0, s>>>(c->prev.p, c->prevCount, off, apply, bad);
            if (c->eventCount)
                kShiftEvents<<<blocks(c->eventCount), kThreads, 0, s>>>(c->events.p, c->eventCount, off, apply, bad);
            HIPCK(hipGetLastError());
            if (!apply) {
                uint32_t flag = 0;
                HIPCK(hipMemcpyAsync(&flag, bad, sizeof(flag), hipMemcpyDeviceToHost, s));
                HIPCK(hipStreamSynchronize(s));
                if (flag) fail(VXHIP_ERR_INVALID, "velox: origin shift overflows world coordinates");
            }
        }
        HIPCK(hipStreamSynchronize(s));
    });
}

namespace {
struct StateHeader {
    uint32_t magic, version, n, prevCount, evPrevCount, eventCount, solveCount, reserved;
};
constexpr uint32_t kStateMagic = 0x54535856u;  // "VXST"

size_t stateBytes(const vxhip_ctx* c) {
    return sizeof(StateHeader) + (size_t)c->n * sizeof(vx_body) + (size_t)c->prevCount * sizeof(vx_contact) +
           (size_t)c->evPrevCount * sizeof(uint64_t) + (size_t)c->eventCount * sizeof(vx_contact_event) +
           (size_t)c->solveCount * sizeof(vx_contact) + (c->prevExact ? 0 : (size_t)c->prevCount * sizeof(uint32_t));
}
// Version 2: the cache's order flag in `reserved` (bit 0: in the reference's
// std::sort order), and, for a canonical cache, each entry's solve position.
constexpr uint32_t kStateVersion = 2u;
}  // namespace

int vxhip_world_state_size(vxhip_ctx* c, size_t* bytes) {
    return guarded([&] {
        if (!c || !c->loaded || !bytes) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        *bytes = stateBytes(c);
    });
}

int vxhip_world_save_state(vxhip_ctx* c, void* out, size_t cap) {
    return guarded([&] {
        if (!c || !c->loaded || !out) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        if (cap < stateBytes(c)) fail(VXHIP_ERR_CAPACITY, "vxhip: state buffer too small");
        hipStream_t s = c->stream;
        uint8_t* p = (uint8_t*)out;
        const StateHeader h{kStateMagic, kStateVersion, c->n, c->prevCount, c->evPrevCount, c->eventCount,
                            c->solveCount, c->prevExact ? 1u : 0u};
        std::memcpy(p, &h, sizeof h);
        p += sizeof h;
        auto down = [&](const void* src, size_t bytes) {
            if (bytes) HIPCK(hipMemcpyAsync(p, src, bytes, hipMemcpyDeviceToHost, s));
            p += bytes;
        };
        down(c->bodies.p, (size_t)c->n * sizeof(vx_body));
        down(c->prev.p, (size_t)c->prevCount * sizeof(vx_contact));
        down(c->evPrevKeys.p, (size_t)c->evPrevCount * sizeof(uint64_t));
        down(c->events.p, (size_t)c->eventCount * sizeof(vx_contact_event));
        down(c->solve.p, (size_t)c->solveCount * sizeof(vx_contact));
        if (!c->prevExact) down(c->prevRank.p, (size_t)c->prevCount * sizeof(uint32_t));
        HIPCK(hipStreamSynchronize(s));
    });
}

int vxhip_world_restore_state(vxhip_ctx* c, const void* in, size_t size) {
    return guarded([&] {
        if (!c || !c->loaded || !in || size < sizeof(StateHeader)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        StateHeader h;
        std::memcpy(&h, in, sizeof h);
        if (h.magic != kStateMagic || h.version != kStateVersion)
            fail(VXHIP_ERR_INVALID, "vxhip: not a saved world state");
        const bool exact = (h.reserved & 1u) != 0;
        if (h.n != c->n) fail(VXHIP_ERR_INVALID, "vxhip: saved state belongs to a world of another size");
        const size_t want = sizeof h + (size_t)h.n * sizeof(vx_body) + (size_t)h.prevCount * sizeof(vx_contact) +
                            (size_t)h.evPrevCount * sizeof(uint64_t) + (size_t)h.eventCount * sizeof(vx_contact_event) +
                            (size_t)h.solveCount * sizeof(vx_contact) +
                            (exact ? 0 : (size_t)h.prevCount * sizeof(uint32_t));
        if (size != want) fail(VXHIP_ERR_INVALID, "vxhip: saved state has the wrong size");
        const uint8_t* p = (const uint8_t*)in + sizeof h;
        std::vector<vx_body> bodies(h.n);
        if (h.n) std::memcpy(bodies.data(), p, (size_t)h.n * sizeof(vx_body));
        p += (size_t)h.n * sizeof(vx_body);
        setBodies(c, bodies.data(), h.n);
        hipStream_t s = c->stream;
        auto up = [&](auto& buf, size_t count, size_t elem) {
            buf.reserve(count ? count : 1);
            if (count) HIPCK(hipMemcpyAsync(buf.p, p, count * elem, hipMemcpyHostToDevice, s));
            p += count * elem;
        };
        up(c->prev, h.prevCount, sizeof(vx_contact));
        up(c->evPrevKeys, h.evPrevCount, sizeof(uint64_t));
        up(c->events, h.eventCount, sizeof(vx_contact_event));
        up(c->solve, h.solveCount, sizeof(vx_contact));
        if (!exact) up(c->prevRank, h.prevCount, sizeof(uint32_t));
        c->prevExact = exact;
        c->prevCount = h.prevCount;
        c->evPrevCount = h.evPrevCount;
        c->eventCount = h.eventCount;
        c->solveCount = h.solveCount;
        HIPCK(hipStreamSynchronize(s));
    });
}

int vxhip_world_set_sleep(vxhip_ctx* c, const vx_sleep_config* config) {
    return guarded([&] {
        if (!c) fail(VXHIP_ERR_INVALID, "vxhip: null context");
        c->sleepStats = vx_sleep_stats{};
        if (!config) {
            c->sleepOn = false;
            return;
        }
        // A 1,001-body world with sleeping on faulted the GPU once (cause not
        // yet found; worlds of a few dozen bodies are verified bit for bit),
        // so sleeping is refused above kSleepMaxBodies until it is.
        if (c->n > kSleepMaxBodies)
            fail(VXHIP_ERR_CAPACITY, "vxhip: device sleeping is limited to 256 bodies in this build");
        vx_sleep_config v = *config;  // SleepConfig::validate (sleep.h:143-159)
        if (v.linearVelocityThreshold < 0.0f) v.linearVelocityThreshold = 0.0f;
        if (v.angularVelocityThreshold < 0.0f) v.angularVelocityThreshold = 0.0f;
        if (v.accelerationThreshold < 0.0f) v.accelerationThreshold = 0.0f;
        if (v.timeToSleep < 0.0f) v.timeToSleep = 0.0f;
        if (v.timeToDrowsy < 0.0f) v.timeToDrowsy = 0.0f;
        if (v.timeToSleep > 0.0f && v.timeToDrowsy >= v.timeToSleep) v.timeToDrowsy = v.timeToSleep * 0.5f;
        if (v.timeToSleep == 0.0f) v.timeToDrowsy = 0.0f;
        if (v.drowsySimulationRate <= 0.0f) v.drowsySimulationRate = 0.01f;
        if (v.drowsySimulationRate > 1.0f) v.drowsySimulationRate = 1.0f;
        if (v.contactStabilityThreshold < 0.0f) v.contactStabilityThreshold = 0.0f;
        if (v.contactStabilityFrames < 0) v.contactStabilityFrames = 0;
        if (v.contactStabilityFrames > 65535) v.contactStabilityFrames = 65535;
        c->sleepCfg = v;
        vxk::SleepParams& p = c->sleepParams;
        p.linTolSq = v.linearVelocityThreshold * v.linearVelocityThreshold;
        p.angTolSq = v.angularVelocityThreshold * v.angularVelocityThreshold;
        p.accTolSq = v.accelerationThreshold * v.accelerationThreshold;
        p.stabilityThreshold = v.contactStabilityThreshold;
        p.timeToSleep = v.timeToSleep;
        p.timeToDrowsy = v.enableGradualSleep ? v.timeToDrowsy : v.timeToSleep;
        p.stabilityFrames = (uint16_t)v.contactStabilityFrames;
        p.gradual = v.enableGradualSleep ? 1u : 0u;
        p.stability = v.enableContactStability ? 1u : 0u;
        c->slStability.reserve(c->n ? c->n : 1);
        HIPCK(hipMemsetAsync(c->slStability.p, 0, (size_t)(c->n ? c->n : 1) * sizeof(uint16_t), c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        c->sleepOn = true;
    });
}

int vxhip_world_sleep_stats(vxhip_ctx* c, vx_sleep_stats* out) {
    return guarded([&] {
        if (!c || !out) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        *out = c->sleepStats;
    });
}

int vxhip_world_stats(vxhip_ctx* c, vxhip_stats* out) {
    return guarded([&] {
        if (!c || !out) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        *out = c->stats;
    });
}

uint32_t vxhip_last_velocity_iterations(const vxhip_ctx* c) { return c ? c->lastIterations : 0; }

int vxhip_world_profile(vxhip_ctx* c, int enable) {
    return guarded([&] {
        if (!c) fail(VXHIP_ERR_INVALID, "vxhip: null context");
        if (enable && c->stageEv.empty()) {
            c->stageEv.resize(sCount);
            for (hipEvent_t& e : c->stageEv) HIPCK(hipEventCreate(&e));
        }
        c->profile = enable != 0;
        c->timings = vxhip_timings{};
        c->launchUsed = 0;
        c->launchContacts.clear();
        c->launchKind.clear();
    });
}

int vxhip_world_timings(vxhip_ctx* c, vxhip_timings* out) {
    return guarded([&] {
        if (!c || !out) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        *out = c->timings;
    });
}

int vxhip_world_record(vxhip_ctx* c, int enable) {
    return guarded([&] {
        if (!c) fail(VXHIP_ERR_INVALID, "vxhip: null context");
        c->recordOn = enable != 0;
        c->record.m = 0;
    });
}

int vxhip_world_download_record(vxhip_ctx* c, vx_contact* contacts, uint32_t* match, uint32_t cap, uint32_t* count) {
    return guarded([&] {
        if (!c || !count) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        if (!c->recordOn) fail(VXHIP_ERR_INVALID, "vxhip: recording is off");
        const uint32_t m = c->record.m;
        *count = m;
        if (!m || cap < m) return;
        if (contacts)
            HIPCK(hipMemcpyAsync(contacts, c->record.cs.p, (size_t)m * sizeof(vx_contact), hipMemcpyDeviceToHost, c->stream));
        if (match)
            HIPCK(hipMemcpyAsync(match, c->record.match.p, (size_t)m * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int vxhip_diff_forward(vxhip_ctx* c, float dt, uint32_t steps, uint32_t every, const float* targets, float* loss) {
    return guarded([&] {
        if (!c || !targets || steps == 0 || every == 0) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        diffValidate(c);
        hipStream_t s = c->stream;
        const uint32_t W = c->diffW, n = c->n;
        c->diffReady = false;
        c->diffSteps = steps;
        c->diffEvery = every;
        c->diffDt = dt;
        const uint32_t segs = (steps + every - 1) / every;
        c->diffCps.resize(segs);
        for (auto& cp : c->diffCps)
            if (!cp) cp.reset(new DiffCheckpoint());
        for (uint32_t t = 0; t < steps; ++t) {
            if (t % every == 0) diffSave(c, *c->diffCps[t / every]);
            worldStep(c, dt);
        }
        c->diffFinal.reserve(n);
        HIPCK(hipMemcpyAsync(c->diffFinal.p, c->bodies.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToDevice, s));
        c->diffTarget.reserve(3 * W); c->diffLoss.reserve(W);
        HIPCK(hipMemcpyAsync(c->diffTarget.p, targets, 3 * (size_t)W * sizeof(float), hipMemcpyHostToDevice, s));
        kDiffLoss<<<blocks(W), kThreads, 0, s>>>(c->diffFinal.p, c->diffBodyStart.p, W, c->diffTarget.p,
                                                 c->diffLoss.p, nullptr, nullptr, nullptr, nullptr);
        HIPCK(hipGetLastError());
        if (loss) HIPCK(hipMemcpyAsync(loss, c->diffLoss.p, (size_t)W * sizeof(float), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        c->diffReady = true;
    });
}

int vxhip_diff_backward(vxhip_ctx* c, const float* lossWeight, float* gV, float* gW, float* gFriction, float* gMass,
                        vxhip_diff_stats* stats) {
    return guarded([&] {
        if (!c || !lossWeight) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        if (!c->diffReady) fail(VXHIP_ERR_INVALID, "vxhip: no differentiable rollout to take gradients of");
        hipStream_t s = c->stream;
        const uint32_t n = c->n, W = c->diffW, T = c->diffSteps, E = c->diffEvery;
        for (Buf<float4>* b : {&c->diffAdjX, &c->diffAdjQ, &c->diffAdjV, &c->diffAdjW, &c->diffAdjParam})
            zeroF4(c, *b, n);
        c->diffAdjMass.reserve(n);
        kFillF<<<blocks(n), kThreads, 0, s>>>(c->diffAdjMass.p, n);
        c->diffOverflow.reserve(2);
        fill32(c, c->diffOverflow.p, 2, 0u);
        c->diffLossBar.reserve(W);
        HIPCK(hipMemcpyAsync(c->diffLossBar.p, lossWeight, (size_t)W * sizeof(float), hipMemcpyHostToDevice, s));
        kDiffLoss<<<blocks(W), kThreads, 0, s>>>(c->diffFinal.p, c->diffBodyStart.p, W, c->diffTarget.p, nullptr,
                                                 nullptr, c->diffLossBar.p, c->diffAdjX.p, c->diffAdjMass.p);
        uint32_t ccd = 0, maxContacts = 0;
        bool last = true;
        const uint32_t segs = (T + E - 1) / E;
        for (uint32_t seg = segs; seg-- > 0;) {
            const uint32_t t0 = seg * E, t1 = std::min(T, t0 + E);
            diffRestore(c, *c->diffCps[seg]);
            c->diffSeg.resize(t1 - t0);
            for (uint32_t j = 0; j < t1 - t0; ++j) {
                if (!c->diffSeg[j]) c->diffSeg[j].reset(new DiffStepRec());
                c->diffCap = c->diffSeg[j].get();
                try {
                    worldStep(c, c->diffDt);
                } catch (...) {
                    c->diffCap = nullptr;
                    throw;
                }
                c->diffCap = nullptr;
                ccd += c->diffSeg[j]->ccdRescued;
                maxContacts = std::max(maxContacts, c->diffSeg[j]->m);
            }
            // The state after the segment (the replay's reference): the next
            // checkpoint's bodies, or the final bodies.
            const vx_body* after = seg + 1 < segs ? c->diffCps[seg + 1]->bodies.p : c->diffFinal.p;
            for (uint32_t j = t1 - t0; j-- > 0;) {
                const DiffStepRec& r = *c->diffSeg[j];
                if (last) {
                    zeroF4(c, c->diffLamBar, r.m);
                    last = false;
                } else {
                    std::swap(c->diffLamBar.p, c->diffLamPrev.p);
                    std::swap(c->diffLamBar.cap, c->diffLamPrev.cap);
                    c->diffLamBar.reserve(r.m ? r.m : 1);
                }
                diffBackStep(c, r, j + 1 < t1 - t0 ? c->diffSeg[j + 1]->start.p : after);
            }
        }
        c->diffG.reserve(8 * (size_t)n);
        float* g = c->diffG.p;
        kDiffParams<<<blocks(n), kThreads, 0, s>>>(c->bodies.p, n, c->diffAdjV.p, c->diffAdjW.p, c->diffAdjParam.p,
                                                   c->diffAdjMass.p, g, g + 3 * (size_t)n, g + 6 * (size_t)n,
                                                   g + 7 * (size_t)n);
        HIPCK(hipGetLastError());
        if (gV) HIPCK(hipMemcpyAsync(gV, g, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
        if (gW) HIPCK(hipMemcpyAsync(gW, g + 3 * (size_t)n, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
        if (gFriction)
            HIPCK(hipMemcpyAsync(gFriction, g + 6 * (size_t)n, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
        if (gMass) HIPCK(hipMemcpyAsync(gMass, g + 7 * (size_t)n, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
        uint32_t over[2] = {0, 0};
        HIPCK(hipMemcpyAsync(over, c->diffOverflow.p, sizeof over, hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        // The replays leave the world at the end of the first segment: put it
        // back at the rollout's end.
        HIPCK(hipMemcpyAsync(c->bodies.p, c->diffFinal.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToDevice, s));
        c->prevCount = 0;
        c->prevExact = false;
        HIPCK(hipStreamSynchronize(s));
        if (stats) {
            stats->steps = T;
            stats->checkpoints = segs;
            stats->ccdRescued = ccd;
            stats->replayMismatch = over[1];
            stats->maxContacts = maxContacts;
            stats->worlds = W;
        }
        if (over[0]) fail(VXHIP_ERR_CAPACITY, "vxhip: a world has more colour runs than the adjoint kernels hold");
    });
}

int vxhip_invalidate_caches(vxhip_ctx* c) {
    return guarded([&] {
        if (!c) fail(VXHIP_ERR_INVALID, "vxhip: null context");
        c->solveCount = 0;
        c->prevCount = 0;
        c->prevExact = false;
        for (size_t& s : c->soupSig) s = (size_t)-1;
    });
}


// --------------------------------------------------- Backend equivalents --
// These take host records every call, like the reference's device backend
// (backend_cuda.cu:393-768); they share the context's device buffers with the
// resident world (calling them replaces the resident bodies).

int vxhip_integrate(vxhip_ctx* c, vx_body* bodies, uint32_t n, const float gravity[3], float dt) {
    return guarded([&] {
        injectFailureForTesting();
        if (!c || (!bodies && n) || !gravity) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        if (n == 0) return;
        detachWorld(c);
        setBodies(c, bodies, n);
        kIntegrate<<<blocks(n), kThreads, 0, c->stream>>>(c->bodies.p, n, V3{gravity[0], gravity[1], gravity[2]}, dt);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(bodies, c->bodies.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

static void unpackPairs(uint64_t* keys, uint32_t np, int nb) {
    for (uint32_t p = 0; p < np; ++p) {
        uint64_t k = keys[p];
        keys[p] = ((k >> nb) << 32) | (k & ((1ull << nb) - 1));
    }
}

int vxhip_probe_pairs(vxhip_ctx* c, const vx_body* bodies, uint32_t n, const vx_soup* soup, float dt, uint64_t* out,
                      uint32_t cap, uint32_t* count) {
    return guarded([&] {
        if (!c || !count || (!bodies && n)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        validateScene(bodies, n, soup);
        *count = 0;
        if (n < 2) return;
        uploadSoup(c, soup);
        detachWorld(c);
        setBodies(c, bodies, n);
        const uint32_t np = broadPhase(c, dt);
        *count = np;
        if (np > cap) fail(VXHIP_ERR_CAPACITY, "vxhip: pair buffer too small");
        if (np) HIPCK(hipMemcpyAsync(out, c->pairs.p, (size_t)np * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        unpackPairs(out, np, bodyBits(n));
    });
}

int vxhip_find_contacts(vxhip_ctx* c, const vx_body* bodies, uint32_t n, const vx_soup* soup, float dt,
                        vx_contact* out, uint32_t cap, uint32_t* count) {
    return guarded([&] {
        injectFailureForTesting();
        if (!c || !count || (!bodies && n)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        validateScene(bodies, n, soup);
        *count = 0;
        if (n < 2) return;
        uploadSoup(c, soup);
        detachWorld(c);
        setBodies(c, bodies, n);
        const uint32_t np = broadPhase(c, dt);
        const uint32_t m = narrowPhase(c, np, dt);
        *count = m;
        if (m > cap) fail(VXHIP_ERR_CAPACITY, "vxhip: contact buffer too small");
        if (m) HIPCK(hipMemcpyAsync(out, c->contacts.p, (size_t)m * sizeof(vx_contact), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int vxhip_probe_color(vxhip_ctx* c, const vx_body* bodies, uint32_t n, const vx_contact* contacts, uint32_t m,
                      uint32_t* colors, uint32_t* numColors) {
    return guarded([&] {
        if (!c || !numColors || (!bodies && n) || (!contacts && m)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        detachWorld(c);
        setBodies(c, bodies, n);
        c->contacts.reserve(m ? m : 1);
        if (m) HIPCK(hipMemcpyAsync(c->contacts.p, contacts, (size_t)m * sizeof(vx_contact), hipMemcpyHostToDevice, c->stream));
        const uint32_t K = colorContacts(c, c->contacts.p, m);
        *numColors = K;
        if (m) HIPCK(hipMemcpyAsync(colors, c->color.p, (size_t)m * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int vxhip_probe_warm_start(vxhip_ctx* c, vx_contact* cur, uint32_t m, const vx_contact* prev, uint32_t mp) {
    return guarded([&] {
        if (!c || (!cur && m) || (!prev && mp)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        if (m == 0) return;
        c->contacts.reserve(m);
        c->prev.reserve(mp ? mp : 1);
        HIPCK(hipMemcpyAsync(c->contacts.p, cur, (size_t)m * sizeof(vx_contact), hipMemcpyHostToDevice, c->stream));
        if (mp) HIPCK(hipMemcpyAsync(c->prev.p, prev, (size_t)mp * sizeof(vx_contact), hipMemcpyHostToDevice, c->stream));
        if (mp) kWarmMatch<<<blocks(m), kThreads, 0, c->stream>>>(c->contacts.p, m, c->prev.p, mp);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(cur, c->contacts.p, (size_t)m * sizeof(vx_contact), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        c->prevCount = 0;
        c->prevExact = false;
    });
}

int vxhip_std_sort_noalloc(vxhip_ctx* c, uint64_t* keys, uint32_t* vals, uint32_t n, const uint32_t* segStart,
                           uint32_t segments) {
    return guarded([&] {
        if (!c || (n && (!keys || !vals)) || (segStart && segments == 0)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        if (n > kNoAllocMaxKeys || (segStart && segments > kNoAllocMaxSegments))
            fail(VXHIP_ERR_CAPACITY, "vxhip: the allocation-free sort takes at most 262,144 keys and 4,096 segments");
        const uint32_t W = segStart ? segments : 1u;
        if (vxhost::stdSortTempBytes(n, W) + 256 > kNoAllocScratchBytes)
            fail(VXHIP_ERR_CAPACITY, "vxhip: the allocation-free sort's working memory exceeds its 4 MiB array");
        int dev = 0;
        HIPCK(hipGetDevice(&dev));
        if (dev != c->device) fail(VXHIP_ERR_INVALID, "vxhip: the context belongs to another device");
        uint8_t* base = scratchOn(dev);
        uint32_t* bad = reinterpret_cast<uint32_t*>(base);
        if (n <= 1) return;
        hipStream_t s = c->stream;
        kFill32<<<1, 64, 0, s>>>(bad, 1, 0u);
        HIPCK(vxhost::stdSort(keys, vals, n, segStart, W, base + 256, s, segStart ? bad : nullptr));
        HIPCK(hipMemcpyAsync(c->hostCounters + kNoAllocFlagSlot, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        if (c->hostCounters[kNoAllocFlagSlot])
            fail(VXHIP_ERR_INVALID, "vxhip: segment offsets must rise from 0 to n without falling");
    });
}

int vxhip_probe_radix_sort(vxhip_ctx* c, uint64_t* keys, uint32_t* vals, uint32_t n, int bits) {
    return guarded([&] {
        if (!c || (n && (!keys || !vals)) || bits < 0 || bits > 64) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        if (n <= 1) return;
        hipStream_t s = c->stream;
        c->exKeys.reserve(n); c->exVals.reserve(n); c->exWorld.reserve(n); c->exIdx.reserve(n);
        HIPCK(hipMemcpyAsync(c->exKeys.p, keys, (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(c->exVals.p, vals, (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        HIPCK(vxhost::radixSort(c->exKeys.p, c->exVals.p, c->exWorld.p, c->exIdx.p, n, bits,
                                tempFor(c, vxhost::radixTempBytes(n)), s));
        HIPCK(hipMemcpyAsync(keys, c->exKeys.p, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipMemcpyAsync(vals, c->exVals.p, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
    });
}

int vxhip_probe_std_sort(vxhip_ctx* c, uint64_t* keys, uint32_t* vals, uint32_t n, const uint32_t* segStart,
                         uint32_t segments) {
    return guarded([&] {
        if (!c || (n && (!keys || !vals)) || (segStart && segments == 0)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        if (segStart) {
            for (uint32_t w = 0; w < segments; ++w)
                if (segStart[w] > segStart[w + 1]) fail(VXHIP_ERR_INVALID, "vxhip: segment starts must not decrease");
            if (segStart[0] != 0 || segStart[segments] != n) fail(VXHIP_ERR_INVALID, "vxhip: segments must cover [0, n)");
        }
        if (n <= 1) return;
        hipStream_t s = c->stream;
        const uint32_t W = segStart ? segments : 1u;
        c->exKeys.reserve(n); c->exVals.reserve(n); c->exStart.reserve(W + 1);
        HIPCK(hipMemcpyAsync(c->exKeys.p, keys, (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(c->exVals.p, vals, (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        if (segStart)
            HIPCK(hipMemcpyAsync(c->exStart.p, segStart, (size_t)(W + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        HIPCK(vxhost::stdSort(c->exKeys.p, c->exVals.p, n, segStart ? c->exStart.p : nullptr, W,
                              tempFor(c, vxhost::stdSortTempBytes(n, W)), s));
        HIPCK(hipMemcpyAsync(keys, c->exKeys.p, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipMemcpyAsync(vals, c->exVals.p, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
    });
}

static bool hostCanonicalLess(const vx_contact& x, const vx_contact& y) {  // backend_cuda.cu:587-595
    uint64_t kx = (uint64_t)x.a << 32 | x.b, ky = (uint64_t)y.a << 32 | y.b;
    if (kx != ky) return kx < ky;
    if (x.point.x != y.point.x) return x.point.x < y.point.x;
    if (x.point.y != y.point.y) return x.point.y < y.point.y;
    return x.point.z < y.point.z;
}

// First-substep work of solveVelocities: canonical order (std::sort with
// the reference's comparator, so contacts that compare equal end where
// libstdc++ leaves them there), colour, colour order.  Leaves the
// colour-ordered contacts in c->solve.
static void prepareSolve(vxhip_ctx* c, vx_contact* contacts, uint32_t m) {
    std::sort(contacts, contacts + m, hostCanonicalLess);
    c->contacts.reserve(m);
    HIPCK(hipMemcpyAsync(c->contacts.p, contacts, (size_t)m * sizeof(vx_contact), hipMemcpyHostToDevice, c->stream));
    const uint32_t K = colorContacts(c, c->contacts.p, m);
    colorOrder(c, m, K);
    c->solve.reserve(m);
    kGather<<<blocks(m), kThreads, 0, c->stream>>>(c->contacts.p, c->order.p, m, c->solve.p);
    c->numColors = K;
    c->solveCount = m;
}

int vxhip_solve_velocities(vxhip_ctx* c, vx_body* bodies, uint32_t n, vx_contact* contacts, uint32_t m, float dt,
                           int warmStart, const vx_solver_options* options) {
    return guarded([&] {
        injectFailureForTesting();
        if (!c || !options || (!bodies && n) || (!contacts && m)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        c->lastIterations = 0;
        if (m == 0) return;
        if (!warmStart && c->solveCount != m) fail(VXHIP_ERR_INVALID, "vxhip: solve without a warm-started contact set");
        detachWorld(c);
        setBodies(c, bodies, n);
        if (warmStart) {
            prepareSolve(c, contacts, m);
            // The caller's contacts come back in colour order with the
            // impulses they came in with (the reference permutes its host
            // copy and solves on the device one; fetchImpulses reads those).
            HIPCK(hipMemcpyAsync(contacts, c->solve.p, (size_t)m * sizeof(vx_contact), hipMemcpyDeviceToHost, c->stream));
            // Later substeps replay the sweeps recorded here, with this
            // call's step and options (the reference captures them in a
            // graph at the warm-started call and replays it).
            c->solveH = dt;
            c->solveOptions = *options;
        }
        substepSolve(c, c->solve.p, c->numColors, c->solveH, c->solveOptions, warmStart != 0, true);
        // Only the warm-started call reports its sweeps (backend_cuda.cu:573, 665).
        if (warmStart) c->lastIterations = (uint32_t)(2 * c->solveOptions.velocityIterations);
        HIPCK(hipMemcpyAsync(bodies, c->bodies.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

int vxhip_advance_substeps(vxhip_ctx* c, vx_body* bodies, uint32_t n, vx_contact* contacts, uint32_t m,
                           vx_joint* joints, uint32_t nj, const float gravity[3], float h, int substeps,
                           const vx_solver_options* options, int* advanced) {
    return guarded([&] {
        injectFailureForTesting();
        if (!c || !options || !advanced || !gravity || (!bodies && n) || (!contacts && m) || (!joints && nj))
            fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        *advanced = 0;
        // Fewer than 64 joints or an unknown type stay on the host (backend_cuda.cu:696-701).
        if (nj && nj < kMinDeviceJoints) return;
        for (uint32_t k = 0; k < nj; ++k) {
            if (joints[k].type > VX_JOINT_GEAR) return;
            if (joints[k].a >= n || joints[k].b >= n) fail(VXHIP_ERR_INVALID, "vxhip: joint body index out of range");
        }
        // A joint between an asleep and an awake body wakes both, in joint
        // order, before the policy check (backend_cuda.cu:703-707).
        for (uint32_t k = 0; k < nj; ++k) {
            vx_body& a = bodies[joints[k].a];
            vx_body& b = bodies[joints[k].b];
            if (a.asleep == b.asleep) continue;
            a.asleep = b.asleep = 0;
            a.sleepTimer = b.sleepTimer = 0.0f;
        }
        if (options->iterationPolicy == VX_ITERATION_ADAPTIVE) return;  // backend_cuda.cu:714-715
        detachWorld(c);
        setBodies(c, bodies, n);
        const V3 g{gravity[0], gravity[1], gravity[2]};
        if (m) prepareSolve(c, contacts, m);
        else c->solveCount = 0;
        // The joints of this call replace any a loaded world holds; the
        // context's body buffer is the call's too.
        c->hostJoints.assign(joints, joints + nj);
        prepareJoints(c);
        const int S = substeps > 0 ? substeps : 1;
        for (int s = 0; s < S; ++s) {
            if (s > 0) kIntegrate<<<blocks(n), kThreads, 0, c->stream>>>(c->bodies.p, n, g, h);
            if (m) substepSolve(c, c->solve.p, c->numColors, h, *options, s == 0, s == S - 1);
            substepJoints(c, h, s == 0);
            kTransforms<<<blocks(n), kThreads, 0, c->stream>>>(c->bodies.p, n, h);
        }
        HIPCK(hipGetLastError());
        c->lastIterations = (uint32_t)(2 * options->velocityIterations * S);
        HIPCK(hipMemcpyAsync(bodies, c->bodies.p, (size_t)n * sizeof(vx_body), hipMemcpyDeviceToHost, c->stream));
        if (m) HIPCK(hipMemcpyAsync(contacts, c->solve.p, (size_t)m * sizeof(vx_contact), hipMemcpyDeviceToHost, c->stream));
        if (nj) HIPCK(hipMemcpyAsync(joints, c->joints.p, (size_t)nj * sizeof(vx_joint), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        *advanced = 1;
    });
}

int vxhip_fetch_impulses(vxhip_ctx* c, vx_contact* contacts, uint32_t m) {
    return guarded([&] {
        if (!c || (!contacts && m)) fail(VXHIP_ERR_INVALID, "vxhip: invalid arguments");
        if (c->solveCount == 0 || m != c->solveCount) return;  // backend_cuda.cu:763-768
        HIPCK(hipMemcpyAsync(contacts, c->solve.p, (size_t)m * sizeof(vx_contact), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
