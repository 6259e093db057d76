/*
 * vx_records.h -- plain-C record layouts shared by the HIP device step and its
 * callers.  Every record is byte-compatible with the Velox record it mirrors, so
 * a host that owns `std::vector<velox::Body>` / `std::vector<velox::Contact>`
 * can hand its buffers to the vxhip_* entry points without conversion.
 *
 *   vx_body            <-> velox::Body            include/velox/body.h:128-171  (320 B, align 64)
 *   vx_ccd_tuning      <-> velox::BodyCcdTuning   include/velox/ccd.h:89-96     (52 B)
 *   vx_contact         <-> velox::Contact         include/velox/backend.h:63-83 (128 B, align 64)
 *   vx_mesh            <-> velox::Mesh            include/velox/body.h:354-362
 *   vx_bvh_node        <-> velox::BvhNode         include/velox/body.h:370-374
 *   vx_compound_child  <-> velox::CompoundChild   include/velox/body.h:105-115
 *   vx_solver_options  <-> velox::SolverOptions   include/velox/solver.h:23-35 (24 B)
 *
 * Enumerations keep the reference's numeric values (they are part of the
 * record format).
 */
#ifndef VX_RECORDS_H
#define VX_RECORDS_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z; } vx_vec3;
typedef struct { float x, y, z, w; } vx_quat;

/* ShapeType (body.h:64-67) */
enum {
    VX_SHAPE_SPHERE = 0, VX_SHAPE_PLANE = 1, VX_SHAPE_BOX = 2, VX_SHAPE_CAPSULE = 3,
    VX_SHAPE_MESH = 4, VX_SHAPE_HULL = 5, VX_SHAPE_COMPOUND = 6, VX_SHAPE_CYLINDER = 7,
    VX_SHAPE_CONE = 8, VX_SHAPE_ROUNDED_BOX = 9, VX_SHAPE_ELLIPSOID = 10
};
/* MotionType (body.h:70-74) */
enum { VX_MOTION_STATIC = 0, VX_MOTION_KINEMATIC = 1, VX_MOTION_DYNAMIC = 2 };
/* MotionQuality (ccd.h:13) */
enum { VX_QUALITY_LOW = 0, VX_QUALITY_MEDIUM = 1, VX_QUALITY_HIGH = 2, VX_QUALITY_LOCKED = 3 };
/* GeometryQuality (body.h:82) */
enum { VX_GEOM_NORMAL = 0, VX_GEOM_ROBUST = 1, VX_GEOM_PARANOID = 2 };
/* MaterialCombineMode (body.h:89-91) */
enum { VX_COMBINE_AVERAGE = 0, VX_COMBINE_GEOMETRIC_MEAN = 1, VX_COMBINE_MINIMUM = 2,
       VX_COMBINE_MULTIPLY = 3, VX_COMBINE_MAXIMUM = 4 };
/* FrictionModel / IterationPolicy (solver.h:11-19) */
enum { VX_FRICTION_TWO_AXIS = 0, VX_FRICTION_CONE_BLOCK = 1 };
enum { VX_ITERATION_FIXED = 0, VX_ITERATION_ADAPTIVE = 1 };

/* CcdQualitySettings (ccd.h:29-41), 32 B */
typedef struct {
    uint8_t quality; uint8_t _pad0[3];
    uint32_t maxAdvancementIterations;
    uint32_t maxRootFindIterations;
    float distanceTolerance;
    float timeTolerance;
    uint8_t enableRotationalCcd, enableCompoundCcd, useBrentRootFinder, _pad1;
    uint32_t meshBvhMaxDepth;
    float speculativeMargin;
} vx_ccd_settings;

/* BodyCcdTuning (ccd.h:89-96), 52 B */
typedef struct {
    uint8_t quality; uint8_t _pad0[3];
    float collisionMargin;
    float speculativeDistance;
    uint8_t enableContinuous; uint8_t _pad1[3];
    float minVelocityForCCD;
    vx_ccd_settings ccdSettings;
} vx_ccd_tuning;

/* Body (body.h:128-171), 320 B */
typedef struct {
    vx_vec3 position;
    vx_quat orientation;
    vx_vec3 velocity;
    vx_vec3 angularVelocity;
    vx_vec3 force;
    vx_vec3 torque;
    float invMass;
    vx_vec3 invInertia;
    vx_quat inertiaOrientation;
    float restitution;
    float friction;
    vx_vec3 frictionScale;
    float rollingFriction;
    float spinningFriction;
    uint8_t frictionCombine;
    uint8_t restitutionCombine;
    uint8_t _pad0[2];
    float linearDamping;
    float angularDamping;
    float gravityScale;
    uint8_t motionType; uint8_t _pad1[3];
    vx_ccd_tuning ccdTuning;
    uint8_t geometryQuality; uint8_t _pad2[3];
    uint32_t categoryBits;
    uint32_t maskBits;
    int32_t groupIndex;
    uint8_t sensor;
    uint8_t enableSleep;
    uint8_t fixedRotation;
    uint8_t asleep;
    float sleepTimer;
    uint8_t shape; uint8_t _pad3[3];
    float radius;
    vx_vec3 halfExtents;
    float capsuleHalfHeight;
    vx_vec3 planeNormal;
    float planeOffset;
    uint32_t meshIndex;
    uint32_t hullFirst, hullCount;
    uint32_t hullFaceFirst, hullFaceCount;
    uint32_t compoundFirst, compoundCount;
    uint8_t _tail[20];
} vx_body;

/* Contact (backend.h:63-83), 128 B */
typedef struct {
    uint32_t a, b;
    uint64_t featureKey;
    vx_vec3 normal;
    vx_vec3 point;
    vx_vec3 localAnchorA;
    vx_vec3 localAnchorB;
    float gap;
    float bias0;
    float vn0;
    float restitution;
    float friction1, friction2;
    float rollingFriction, spinningFriction;
    float normalImpulse;
    float tangentImpulse1;
    float tangentImpulse2;
    float rollingImpulse1, rollingImpulse2;
    float spinningImpulse;
    uint8_t _tail[8];
} vx_contact;

/* Mesh (body.h:354-362), 52 B */
typedef struct {
    uint32_t firstVertex, vertexCount;
    uint32_t firstIndex, indexCount;
    uint32_t firstNode, nodeCount;
    uint32_t firstTriRef;
    vx_vec3 aabbMin, aabbMax;
} vx_mesh;

/* BvhNode (body.h:370-374), 32 B */
typedef struct {
    vx_vec3 aabbMin, aabbMax;
    uint32_t leftFirst;
    uint32_t triCount;
} vx_bvh_node;

/* CompoundChild (body.h:105-115), 68 B */
typedef struct {
    uint8_t shape; uint8_t _pad0[3];
    vx_vec3 localPosition;
    vx_quat localOrientation;
    float radius;
    vx_vec3 halfExtents;
    float capsuleHalfHeight;
    uint32_t hullFirst, hullCount;
    uint32_t hullFaceFirst, hullFaceCount;
} vx_compound_child;

/* SolverOptions (solver.h:23-35), 24 B */
typedef struct {
    float impulseThreshold;
    float stackStiffness;
    float stackDamping;
    int32_t velocityIterations;
    int32_t positionIterations;
    uint8_t frictionModel;
    uint8_t iterationPolicy;
    uint8_t enableStackStabilization;
    uint8_t reserved;
} vx_solver_options;

/* ContactEvent (world.h:65-71) with dense body indices in place of BodyId
 * handles (the host maps indices to handles), 44 B.  type: 0 Begin,
 * 1 Persist, 2 End (ContactEventType, world.h:46). */
typedef struct {
    uint32_t a, b;
    float point[3];
    float normal[3];
    float impulse;
    uint32_t type;
    uint32_t sensor;
} vx_contact_event;

/* velox::Joint (include/velox/joint.h:99-186), 420 B, byte for byte. */
typedef struct vx_joint {
    uint8_t type;                 /* JointType (joint.h:48-62) */
    uint8_t pad0[3];
    uint32_t a, b;                /* dense body indices */
    vx_vec3 localAnchorA, localAnchorB;
    vx_vec3 localAxisA, localAxisB;
    vx_vec3 localRefA, localRefB;
    float restLength;
    uint8_t collideConnected, enableSpring, pad1[2];
    float springFrequencyHz, springDampingRatio;
    float breakForce, breakTorque;
    uint8_t enableMotor, pad2[3];
    float motorSpeed, maxMotorTorque, maxMotorForce;
    uint8_t enableLimit, pad3[3];
    float lowerLimit, upperLimit;
    uint8_t enableSwingLimit, pad4[3];
    float swingLimit;
    uint8_t enableTwistLimit, pad5[3];
    float lowerTwistLimit, upperTwistLimit;
    uint8_t linearLimitMask, angularLimitMask, linearMotorMask, angularMotorMask;
    vx_vec3 lowerLinearLimit, upperLinearLimit, lowerAngularLimit, upperAngularLimit;
    vx_vec3 linearMotorSpeed, angularMotorSpeed, maxLinearMotorForce, maxAngularMotorTorque;
    float weldFrequencyHz, weldDampingRatio;
    float suspensionFrequencyHz, suspensionDampingRatio, wheelRadius;
    uint8_t enableSteering, pad6[3];
    float steeringAngle, maxSteeringAngle, maxLength;
    vx_vec3 pulleyAnchorA, pulleyAnchorB;
    float pulleyRatio, gearRatio;
    float motorImpulse, limitImpulse, swingImpulse, twistImpulse, springImpulse;
    vx_vec3 linearMotorImpulse, angularMotorImpulse, linearLimitImpulse, angularLimitImpulse;
    vx_vec3 reactionLinearImpulse, reactionAngularImpulse;
    uint8_t broken, pad7[3];
} vx_joint;

enum {
    VX_JOINT_BALL = 0, VX_JOINT_DISTANCE = 1, VX_JOINT_HINGE = 2, VX_JOINT_CONE_TWIST = 3, VX_JOINT_FIXED = 4,
    VX_JOINT_PRISMATIC = 5, VX_JOINT_SIX_DOF = 6, VX_JOINT_MOTOR = 7, VX_JOINT_WELD = 8, VX_JOINT_WHEEL = 9,
    VX_JOINT_ROPE = 10, VX_JOINT_PULLEY = 11, VX_JOINT_GEAR = 12
};

/* Raw-pointer view of the flat static geometry (MeshSoupView, backend.h:25-34)
 * plus the element counts a device copy needs. */
typedef struct {
    const vx_vec3* vertices;            uint32_t vertexCount;
    const uint32_t* indices;            uint32_t indexCount;
    const vx_mesh* meshes;              uint32_t meshCount;
    const vx_bvh_node* bvhNodes;        uint32_t bvhNodeCount;
    const uint32_t* bvhTriRefs;         uint32_t bvhTriRefCount;
    const vx_vec3* hullPoints;          uint32_t hullPointCount;
    const uint32_t* hullFaceIndices;    uint32_t hullFaceIndexCount;
    const vx_compound_child* compoundChildren; uint32_t compoundChildCount;
} vx_soup;

#ifdef __cplusplus
} /* extern "C" */
static_assert(sizeof(vx_ccd_settings) == 32, "CcdQualitySettings layout");
static_assert(sizeof(vx_ccd_tuning) == 52, "BodyCcdTuning layout");
static_assert(sizeof(vx_body) == 320, "Body layout");
static_assert(sizeof(vx_joint) == 420, "Joint layout");
static_assert(offsetof(vx_joint, restLength) == 84 && offsetof(vx_joint, springFrequencyHz) == 92 &&
              offsetof(vx_joint, enableMotor) == 108 && offsetof(vx_joint, enableLimit) == 124 &&
              offsetof(vx_joint, linearLimitMask) == 156 && offsetof(vx_joint, lowerLinearLimit) == 160 &&
              offsetof(vx_joint, weldFrequencyHz) == 256 && offsetof(vx_joint, enableSteering) == 276 &&
              offsetof(vx_joint, pulleyAnchorA) == 292 && offsetof(vx_joint, motorImpulse) == 324 &&
              offsetof(vx_joint, reactionLinearImpulse) == 392 && offsetof(vx_joint, broken) == 416,
              "Joint field offsets (joint.h:99-186)");
static_assert(offsetof(vx_body, ccdTuning) == 156, "Body layout");
static_assert(offsetof(vx_body, shape) == 232, "Body layout");
static_assert(offsetof(vx_body, compoundCount) == 296, "Body layout");
static_assert(sizeof(vx_contact) == 128, "Contact layout");
static_assert(sizeof(vx_contact_event) == 44, "ContactEvent record");
static_assert(offsetof(vx_contact, spinningImpulse) == 116, "Contact layout");
static_assert(sizeof(vx_mesh) == 52, "Mesh layout");
static_assert(sizeof(vx_bvh_node) == 32, "BvhNode layout");
static_assert(sizeof(vx_compound_child) == 68, "CompoundChild layout");
static_assert(sizeof(vx_solver_options) == 24, "SolverOptions layout");
#endif

#endif /* VX_RECORDS_H */
