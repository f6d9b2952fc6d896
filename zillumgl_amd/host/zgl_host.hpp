// C++ host layer above the C ABI (libzgl.so): the reference's scene ingestion, scene
// packing and Integrator classes, for hosts written in C++ as the reference is.
//
//   * loadScene: Scene::load (src/core/Scene.cpp:58-127) with loadMaterial
//     (src/core/MaterialLoader.cpp:5-71), OBJ/MTL meshes as the reference's assimp flags
//     ask for (src/core/Resource.cpp:37-88) and ModelInstance::modelMatrix
//     (src/core/Model.cpp:62-72) in glm's FP32 operation order;
//   * pack: Scene::createGLContext (src/core/Scene.cpp:133-270) into the buffers
//     zgl_upload_scene takes, plus the camera uniforms of NaivePath.cpp:49-60
//     (Camera::update, src/core/Camera.cpp:149-162) and the per-pixel seed texture;
//   * NaivePathIntegrator, LightPathIntegrator, TriplePathIntegrator and
//     GlobalQueuePathIntegrator: the Integrator interface (src/core/Integrator.h:25-52)
//     with renderOnePass / reset / getFrame / resultScale and the reference's counters.
//
// Errors are thrown as std::runtime_error (std::invalid_argument for malformed input).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/zgl.h"

namespace zgl {
namespace host {

enum MaterialType : int32_t { Lambertian = 0, Principled = 1, MetalWorkflow = 2, Dielectric = 3, ThinDielectric = 4 };

// The 64-byte material record (src/core/Material.h:32-53) with its defaults.
struct Material {
    float baseColor[3] = {1.0f, 1.0f, 1.0f};
    float roughness = 1.0f, subsurface = 0.0f, metallic = 0.0f, specular = 1.0f, specularTint = 1.0f;
    float sheen = 0.0f, sheenTint = 1.0f, clearcoat = 0.0f, clearcoatGloss = 0.0f, ior = 1.5f;
    int32_t type = Lambertian;
    void record(float out[16]) const;
};

// A triangle mesh in world space: 3 floats per vertex position and normal, 2 per texcoord.
struct Mesh {
    std::vector<float> positions, normals, texcoords;
    std::vector<uint32_t> indices;  // 3 per triangle
    size_t vertexCount() const { return positions.size() / 3; }
};

// Camera as the XML scene describes it (Scene.cpp:85-111).
struct CameraDesc {
    double pos[3] = {0.0, 0.0, 0.0};
    double angle[3] = {90.0, 0.0, 0.0};  // yaw, pitch in degrees; roll in radians
    double fov = 45.0, lensRadius = 0.0, focalDist = 1.0;
};

struct SceneDesc {
    std::vector<std::pair<Mesh, Material>> objects;
    std::vector<std::pair<Mesh, std::vector<float>>> lights;  // radiance (r, g, b) per light mesh
    std::vector<float> env;  // envHeight x envWidth x 3, row 0 at the top; empty: black 1 x 1
    int envWidth = 0, envHeight = 0;
    float envRotation = 0.0f;
    CameraDesc camera;
    int width = 256, height = 256;
    int sampler = 1;  // 1 = Sobol
    uint64_t noiseSeed = 0;
};

// Scene::load.  Paths in the file are tried relative to the XML file, then as given.
SceneDesc loadScene(const std::string& xmlPath);

// One mesh per material group of an OBJ file (in order of first use) and each group's
// MTL diffuse colour (1, 1, 1 when the material is not defined).
struct ObjModel {
    std::vector<Mesh> meshes;
    std::vector<std::pair<std::string, std::vector<float>>> materials;
};
ObjModel readObj(const std::string& path);

// glm in FP32: column-major 4 x 4 (m[c * 4 + r]).
struct Mat4 {
    float m[16];
};
Mat4 modelMatrix(const double pos[3], const double scale[3], const double rotDeg[3]);

// Buffers in the reference layout plus the uniforms; buffers() points into them.
struct PackedScene {
    std::vector<float> vertices, normals, texcoords, bounds, materials, lightPower, lightProb, envProb, noise;
    std::vector<int32_t> indices, hitTable, matTexIndices, lightAlias, envAlias;
    std::vector<uint16_t> envHalf;
    int32_t objPrimCount = 0, nLightTriangles = 0, envWidth = 0, envHeight = 0, sampler = 1;
    float lightSum = 0.0f, envSum = 0.0f, envRotation = 0.0f;
    int width = 0, height = 0;
    zgl_camera camera{};
    zgl_scene_buffers buffers() const;
};
// deviceBvh: leave bounds and hit tables out, so that zgl_upload_scene builds the MTBVH
// on the GPU; otherwise the host build (zgl_build_mtbvh).
PackedScene pack(const SceneDesc& desc, bool deviceBvh = false);
zgl_camera cameraUniforms(const CameraDesc& cam, int width, int height);
std::vector<float> noiseTexture(int width, int height, uint64_t seed);

// Uploads scene, camera and seed texture into a context.
void upload(zgl_context* ctx, const PackedScene& ps);

// ----------------------------------------------------------------------- integrators
class Integrator {
public:
    virtual ~Integrator() = default;
    // init(Scene*, w, h): the context holds the uploaded scene; the frame is cleared.
    void init(zgl_context* ctx, int width, int height);
    // n > 1 renders n consecutive passes in one call (the same sums as n calls).
    virtual void renderOnePass(int n = 1) = 0;
    virtual void reset();
    // The running sum the frame texture holds (RGBA32F, W x H).
    std::vector<float> getFrame() const;
    virtual float resultScale() const = 0;
    int curSample() const { return mCurSample; }

protected:
    void render(int32_t integrator, int n, int32_t blocksOnePass = 32, int32_t loopsPerPass = 1);
    zgl_context* mCtx = nullptr;
    int mWidth = 0, mHeight = 0;
    int mCurSample = 0, mFreeCounter = 0;
    int mMaxDepth = 4;
    bool mRussianRoulette = false, mSampleLight = true, mLightEnvUniform = false;
    float mLightPortion = 0.5f;
};

// NaivePath.cpp:77-115; PathIntegParam defaults (Integrator.h:56-66).
class NaivePathIntegrator : public Integrator {
public:
    explicit NaivePathIntegrator(int maxDepth = 4, bool russianRoulette = false, bool sampleLight = true,
                                 bool lightEnvUniform = false, float lightPortion = 0.5f);
    void renderOnePass(int n = 1) override;
    float resultScale() const override { return 1.0f / (float)(mCurSample + 1); }
};

// GlobalQueuePath.cpp:100-176; QueuedPTParam defaults (Integrator.h:211-222).
class GlobalQueuePathIntegrator : public Integrator {
public:
    explicit GlobalQueuePathIntegrator(int maxDepth = 4, bool russianRoulette = false, bool sampleLight = true,
                                       bool lightEnvUniform = false, float lightPortion = 0.5f);
    void renderOnePass(int n = 1) override;
    void reset() override;
    float resultScale() const override { return (float)(1.0 / mSamplePerPixel); }

private:
    double mSamplePerPixel = 1.0;
};

// LightPath.cpp:83-135; LightPathIntegParam defaults.
class LightPathIntegrator : public Integrator {
public:
    static constexpr int WorkgroupSize = 1536;
    explicit LightPathIntegrator(int maxDepth = 4, bool russianRoulette = false, int threadBlocksOnePass = 32);
    void renderOnePass(int n = 1) override;
    void reset() override;
    float resultScale() const override { return (float)(1.0 / mSamplePerPixel); }

private:
    int mBlocks;
    double mSamplePerPixel = 0.0;
};

// TriplePath.cpp:97-156; TriplePathIntegParam defaults.
class TriplePathIntegrator : public Integrator {
public:
    explicit TriplePathIntegrator(int maxDepth = 4, bool russianRoulette = false, int lptBlocksOnePass = 64,
                                  int lptLoopsPerPass = 1);
    void renderOnePass(int n = 1) override;
    void reset() override;
    float resultScale() const override { return (float)(1.0 / mSamplePerPixel); }

private:
    int mBlocks, mLoops;
    double mSamplePerPixel = 1.0;
};

}  // namespace host
}  // namespace zgl
