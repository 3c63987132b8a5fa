// SPlisHSPlasH scene files and the reference's text particle files, host side.
//
// SceneLoader follows Utilities/SceneLoader.{h,cpp} (readScene: the
// Configuration, RigidBodies, FluidModels and FluidBlocks sections with the
// same defaults) and Demos/Common/DemoBase.cpp:539-600 (createFluidBlocks,
// the three block packings).  Rigid bodies are sampled as particles: the
// mesh (OBJ) is scaled, rotated and translated as
// Demos/StaticBoundaryDemo/main.cpp:421-424 does, then sampled on regular
// lattices at spacing 2r (an axis-aligned box exactly on its six faces); the
// reference samples with Poisson disks instead (PoissonDiskSampling.cpp), so
// the points differ while their spacing matches.  Static bodies become wall
// particles; dynamic ones ("isDynamic") rigid bodies of the solver, with the
// mass properties of their closed mesh.  Particle sample files (partio
// .bgeo) are not read.
//
// The text particle files are UnifiedParticleSet::write_to_file /
// load_from_file (SPlisHSPlasH/DFSPH/DFSPH_c_arrays_structure.cpp:424-560):
// a header line "n has_factor_computation velocity_impacted_by_fluid_solver
// is_dynamic_object", then one line "mass x y z vx vy vz" per particle.
#pragma once

#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "dfsph_hip.h"

namespace SPH {

// A JSON value (objects keep their keys sorted; numbers are doubles).
struct JsonValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JsonValue> arr;
    std::map<std::string, JsonValue> obj;
    const JsonValue* get(const std::string& key) const;   // nullptr when absent or not an object
};

JsonValue parse_json(const std::string& text);             // throws std::runtime_error

using Vec3 = std::array<double, 3>;
using Mat3 = std::array<std::array<double, 3>, 3>;

class SceneLoader {
public:
    struct Box {
        Vec3 minX, maxX;
    };
    struct BoundaryData {
        std::string samplesFile, meshFile;
        Vec3 translation{0, 0, 0};
        Mat3 rotation{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
        Vec3 scale{1, 1, 1};
        bool dynamic = false, isWall = false;
        // dynamic bodies (keys of the demos' rigid-body simulator): density in
        // kg/m^3, initial velocity and angular velocity
        double density = 1000.0;
        Vec3 velocity{0, 0, 0}, angularVelocity{0, 0, 0};
    };
    struct FluidData {
        std::string samplesFile;
        Vec3 translation{0, 0, 0};
        Mat3 rotation{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
        double scale = 1.0;
    };
    struct FluidBlock {
        Box box;
        unsigned mode = 0;
        Vec3 initialVelocity{0, 0, 0};
    };
    struct Scene {
        std::vector<BoundaryData> boundaryModels;
        std::vector<FluidData> fluidModels;
        std::vector<FluidBlock> fluidBlocks;
        double timeStepSize = 0.001, particleRadius = 0.025;
        unsigned cflMethod = 1;
        double cflFactor = 0.5, cflMaxTimeStepSize = 0.005;
        unsigned maxIterations = 100;
        double maxError = 0.01;
        unsigned maxIterationsV = 100;
        double maxErrorV = 0.1;
        double viscosity = 0.02, density0 = 1000.0;
        Vec3 gravitation{0.0, -9.81, 0.0};
        bool enableDivergenceSolver = true;
    };
    // Utilities/SceneLoader.cpp:13-300 (the keys this solver uses)
    static void readScene(const std::string& fileName, Scene& scene);
};

// DemoBase::createFluidBlocks (Demos/Common/DemoBase.cpp:539-600)
void createFluidBlocks(const SceneLoader::Scene& scene, std::vector<Vec3>& pos, std::vector<Vec3>& vel);

// Boundary particles of every static rigid body (see the file comment); mesh
// paths are relative to base_path as in the demos.
void sampleRigidBodies(const SceneLoader::Scene& scene, const std::string& base_path, std::vector<Vec3>& pos);

// Dynamic rigid bodies ("isDynamic"): mass, center of mass and inertia of the
// closed mesh at the body's density, particles sampled as for the static
// bodies and given in the principal-axis frame through the center of mass
// (the frame dfsph_add_rigid_body expects), the body's rotation mapping that
// frame to the world.
struct SceneBody {
    dfsph_rigid_body body;
    std::vector<Vec3> particles;
};
std::vector<SceneBody> dynamicRigidBodies(const SceneLoader::Scene& scene, const std::string& base_path);

// Volume integrals of a closed triangle mesh (signed tetrahedra from the
// origin): mass, center of mass, principal moments of inertia and the
// principal axes (columns of a rotation matrix).
void meshMassProperties(const std::vector<Vec3>& v, const std::vector<std::array<uint32_t, 3>>& f, double density,
                        double& mass, Vec3& com, Vec3& moments, Mat3& axes);

// The scene's Configuration as solver parameters (library defaults elsewhere).
void sceneParams(const SceneLoader::Scene& scene, dfsph_params& p);

// UnifiedParticleSet text files (header flags as the reference writes them)
void writeParticleText(const std::string& path, const std::vector<float>& xyz, const std::vector<float>& vel,
                       const std::vector<float>& mass, int has_factor, int impacted, int dynamic);
void readParticleText(const std::string& path, std::vector<float>& xyz, std::vector<float>& vel,
                      std::vector<float>& mass);

}  // namespace SPH
