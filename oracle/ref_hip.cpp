// ORACLE -- test infrastructure only.
//
// The drop-in on the reference side: the HipBackend adapter (a velox::Backend
// whose every virtual is the same-named libvxhip entry point, as a maintainer
// would add it to the reference) compiled against the strict reference World
// and installed in it the way the golden-vector writer installs its CPU
// restatement of the device backend (oracle/ref_harness.cpp, CudaSemantics).
// World::step then runs its own host stages (warm-start matching, CCD net,
// positional correction, events) and calls the device for integration,
// contacts, the substepped solve and fetchImpulses (world.cpp:4014, 4035,
// 4201-4211, 4360).
//
//   ref_hip <outdir> <lattice64|pile>
//
// writes state_<step>.npy at the golden trajectory's steps, for the GPU test
// to compare with tests/golden/<case>/ bit for bit.
#define main refHarnessMain
#include "ref_harness.cpp"
#undef main
#include "../include/vxhip.h"

namespace {

static_assert(sizeof(vx_body) == sizeof(velox::Body), "vx_body mirrors velox::Body");
static_assert(sizeof(vx_contact) == sizeof(velox::Contact), "vx_contact mirrors velox::Contact");
static_assert(sizeof(vx_solver_options) == sizeof(velox::SolverOptions), "vx_solver_options mirrors SolverOptions");

void check(int st) {  // backend.h:95-106, backend_cuda.cu:24-40
    switch (st) {
    case VXHIP_OK: return;
    case VXHIP_ERR_MEMORY: throw BackendFailure(BackendFailureKind::MemoryAllocation, vxhip_last_error());
    case VXHIP_ERR_DEVICE_LOST: throw BackendFailure(BackendFailureKind::DeviceLost, vxhip_last_error());
    case VXHIP_ERR_CAPACITY:
        throw VeloxCapacityExceeded(ErrorCode::BodyCapacityExceeded, vxhip_last_error(), __FILE__, __LINE__, __func__);
    default:
        throw VeloxRuntimeError(ErrorCode::BackendUnavailable, vxhip_last_error(), __FILE__, __LINE__, __func__);
    }
}

vx_soup soupView(const MeshSoup& m) {
    vx_soup s{};
    s.vertices = reinterpret_cast<const vx_vec3*>(m.vertices.data()); s.vertexCount = (uint32_t)m.vertices.size();
    s.indices = m.indices.data(); s.indexCount = (uint32_t)m.indices.size();
    s.meshes = reinterpret_cast<const vx_mesh*>(m.meshes.data()); s.meshCount = (uint32_t)m.meshes.size();
    s.bvhNodes = reinterpret_cast<const vx_bvh_node*>(m.bvhNodes.data()); s.bvhNodeCount = (uint32_t)m.bvhNodes.size();
    s.bvhTriRefs = m.bvhTriRefs.data(); s.bvhTriRefCount = (uint32_t)m.bvhTriRefs.size();
    s.hullPoints = reinterpret_cast<const vx_vec3*>(m.hullPoints.data()); s.hullPointCount = (uint32_t)m.hullPoints.size();
    s.hullFaceIndices = m.hullFaceIndices.data(); s.hullFaceIndexCount = (uint32_t)m.hullFaceIndices.size();
    s.compoundChildren = reinterpret_cast<const vx_compound_child*>(m.compoundChildren.data());
    s.compoundChildCount = (uint32_t)m.compoundChildren.size();
    return s;
}

class HipBackend final : public Backend {
public:
    explicit HipBackend(vxhip_ctx* c) : c_(c) {}
    ~HipBackend() override { vxhip_destroy(c_); }
    const char* name() const override { return "hip"; }
    bool wantsHostPairs() const override { return false; }
    void integrate(std::vector<Body>& b, const Vec3& g, float dt) override {
        const float gv[3] = {g.x, g.y, g.z};
        check(vxhip_integrate(c_, reinterpret_cast<vx_body*>(b.data()), (uint32_t)b.size(), gv, dt));
    }
    void findContacts(const std::vector<Body>& b, const MeshSoup& m, float dt, const std::vector<uint64_t>*,
                      std::vector<Contact>& out) override {
        vx_soup s = soupView(m);
        uint32_t n = 0;
        out.resize(out.capacity() ? out.capacity() : 1024);
        int st;
        while ((st = vxhip_find_contacts(c_, reinterpret_cast<const vx_body*>(b.data()), (uint32_t)b.size(), &s, dt,
                                         reinterpret_cast<vx_contact*>(out.data()), (uint32_t)out.size(), &n)) ==
               VXHIP_ERR_CAPACITY)
            out.resize(n);
        check(st);
        out.resize(n);
    }
    void solveVelocities(std::vector<Body>& b, std::vector<Contact>& cs, float dt, bool warm,
                         const SolverOptions& o) override {
        check(vxhip_solve_velocities(c_, reinterpret_cast<vx_body*>(b.data()), (uint32_t)b.size(),
                                     reinterpret_cast<vx_contact*>(cs.data()), (uint32_t)cs.size(), dt, warm,
                                     reinterpret_cast<const vx_solver_options*>(&o)));
    }
    bool advanceSubsteps(std::vector<Body>& b, std::vector<Contact>& cs, std::vector<Joint>& j, const Vec3& g,
                         float h, int S, const SolverOptions& o) override {
        const float gv[3] = {g.x, g.y, g.z};
        int advanced = 0;
        check(vxhip_advance_substeps(c_, reinterpret_cast<vx_body*>(b.data()), (uint32_t)b.size(),
                                     reinterpret_cast<vx_contact*>(cs.data()), (uint32_t)cs.size(),
                                     reinterpret_cast<vx_joint*>(j.data()), (uint32_t)j.size(), gv, h, S,
                                     reinterpret_cast<const vx_solver_options*>(&o), &advanced));
        ++substepCalls;
        return advanced != 0;
    }
    void fetchImpulses(std::vector<Contact>& cs) override {
        check(vxhip_fetch_impulses(c_, reinterpret_cast<vx_contact*>(cs.data()), (uint32_t)cs.size()));
        ++fetchCalls;
    }
    void invalidateCaches() override { check(vxhip_invalidate_caches(c_)); }
    uint32_t lastVelocityIterations() const override { return vxhip_last_velocity_iterations(c_); }
    int substepCalls = 0, fetchCalls = 0;

private:
    vxhip_ctx* c_;
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <outdir> <lattice64|pile>\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1], name = argv[2];
    vxscene::Scene scene;
    std::vector<int> steps;
    if (name == "lattice64") { scene = vxscene::lattice(4, 4, 4); steps = {1, 2, 10, 60, 120}; }
    else if (name == "pile") { scene = vxscene::pile(160, 5.0f, 6.0f); steps = {1, 5, 30, 90}; }
    else { std::fprintf(stderr, "unknown case %s\n", name.c_str()); return 2; }
    int devices = 0;
    if (vxhip_device_count(&devices) != VXHIP_OK || devices == 0) { std::fprintf(stderr, "no HIP device\n"); return 3; }
    vxhip_ctx* ctx = nullptr;
    check(vxhip_create(&ctx));
    World w(BackendType::Cpu);
    build(w, scene);
    HipBackend* hip = new HipBackend(ctx);
    w.backend_.reset(hip);
    mkdir(out.c_str(), 0755);
    int done = 0;
    for (int target : steps) {
        while (done < target) { w.step(scene.dt); ++done; }
        writeRecords(out + "/state_" + std::to_string(target) + ".npy", states(w.bodies_));
        std::printf("%s step %d: %zu contacts, %d device substep calls, %d impulse fetches\n", name.c_str(), target,
                    w.contacts_.size(), hip->substepCalls, hip->fetchCalls);
        std::fflush(stdout);
    }
    return 0;
}
