// Steps a drape through the C++ classes of include/solidcloth.hpp, used the
// way the reference demo uses ClothPatch and ClothModel
// (Demo/ClothDemo.cpp:100-112, 173-176, 424-427), and writes the final
// positions for tests/test_gpu_facade.py to compare with the same scene
// stepped through the Python binding.
//
// usage: facade_drape n inc offset_y young poisson bend iterations thickness
//        tolerance friction sphere_r ground_y frames out.bin
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "solidcloth.hpp"

using solidcloth::Aabb3;
using solidcloth::Vector3;

int main(int argc, char** argv)
{
    if (argc != 15) {
        std::fprintf(stderr, "usage: %s n inc offset_y young poisson bend iterations thickness tolerance friction "
                             "sphere_r ground_y frames out.bin\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]);
    const float inc = std::strtof(argv[2], nullptr), oy = std::strtof(argv[3], nullptr);
    const float young = std::strtof(argv[4], nullptr), poisson = std::strtof(argv[5], nullptr);
    const float bend = std::strtof(argv[6], nullptr);
    const int iterations = std::atoi(argv[7]);
    const float thickness = std::strtof(argv[8], nullptr), tolerance = std::strtof(argv[9], nullptr);
    const float friction = std::strtof(argv[10], nullptr), radius = std::strtof(argv[11], nullptr);
    const float groundY = std::strtof(argv[12], nullptr);
    const int frames = std::atoi(argv[13]);
    try {
        Physics::CollisionWorld world;
        world.AddCollidable(std::make_shared<Physics::Sphere>(Vector3(0.f, 0.f, 0.f), radius));
        world.AddCollidable(
            std::make_shared<Physics::Walls>(Aabb3(Vector3(-10.f, groundY, -10.f), Vector3(10.f, 10.f, 10.f))));
        Physics::ClothPatch cloth(world);
        Physics::ClothModel& model = cloth.GetModel();
        model.SetUnit(1.f);
        model.SetFEM(true);
        model.SetDihedral(true);
        model.SetShearStiffness(young);
        model.SetPoisson(poisson);
        model.SetBendStiffness(bend);
        model.SetNumIterations(iterations);
        model.SetCollisionFlags(Physics::CF_VERTICES | Physics::CF_WALLS);
        model.SetThickness(thickness);
        model.SetTolerance(tolerance);
        model.SetFriction(friction);
        model.SetDragCoeff(1.f);
        cloth.SetNumSteps(4);
        cloth.Init(n, n, inc, Vector3(0.f, oy, 0.f), true, false);
        // getters read back what the setters stored
        if (model.GetShearStiffness() != young || model.GetPoisson() != poisson || model.GetBendStiffness() != bend ||
            model.GetNumIterations() != iterations || model.GetThickness() != thickness ||
            model.GetTolerance() != tolerance || model.GetFriction() != friction || !model.GetFEM() ||
            !model.GetDihedral() || model.GetCollisionFlags() != (Physics::CF_VERTICES | Physics::CF_WALLS) ||
            cloth.GetNumSteps() != 4) {
            std::fprintf(stderr, "a getter disagrees with its setter\n");
            return 1;
        }
        if (model.GetNumParticles() != (size_t)n * n || model.GetNumTris() != 2 * (size_t)(n - 1) * (n - 1)) {
            std::fprintf(stderr, "unexpected counts\n");
            return 1;
        }
        // a refused call throws
        bool refused = false;
        try {
            cloth.SetNumSteps(0);
        } catch (const solidcloth::Error& e) {
            refused = true;
            std::printf("refused: %s\n", e.what());
        }
        if (!refused || cloth.GetNumSteps() != 4) {
            std::fprintf(stderr, "SetNumSteps(0) was not refused\n");
            return 1;
        }
        // frames: the last one as four single substeps of the PBD model
        for (int f = 0; f + 1 < frames; f++) cloth.Step(1.f / 60);
        Physics::ClothModelPBD& pbd = static_cast<Physics::ClothModelPBD&>(model);
        for (int k = 0; k < 4; k++) pbd.Step(1.f / 240);
        const Physics::SolverStats st = model.GetStats();
        const std::vector<Vector3> pos = model.GetPositions();
        const Physics::Particle p0 = model.GetParticle(0);
        if (p0.pos.x != pos[0].x || p0.pos.y != pos[0].y || p0.pos.z != pos[0].z || p0.invMass <= 0.f) {
            std::fprintf(stderr, "GetParticle disagrees with GetPositions\n");
            return 1;
        }
        FILE* f = std::fopen(argv[14], "wb");
        if (!f || std::fwrite(pos.data(), sizeof(Vector3), pos.size(), f) != pos.size()) {
            std::fprintf(stderr, "cannot write %s\n", argv[14]);
            return 1;
        }
        std::fclose(f);
        std::printf("stepped %d frames: %zu contacts, %d iterations in the last substep\n", frames,
                    model.GetNumContacts(), st.iterations);
    } catch (const solidcloth::Error& e) {
        std::fprintf(stderr, "solidcloth error: %s\n", e.what());
        return 1;
    }
    return 0;
}
