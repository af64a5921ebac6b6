// Incremental scene container behind the C ABI.
//
// Semantics follow the reference's mrt::Scene (core/include/mrt/Scene.hpp,
// core/src/scene/Scene.cpp:13-265): one id counter shared by every object
// kind starting at 1, deep copies of all borrowed data, area-weighted vertex
// normals when none are given, tangents when UVs are given, texture slots
// recycled on removal, and dirty bits consumed by the renderer at commit.
#pragma once

#include "mini_raytrace.h"

#include <cstdint>
#include <map>
#include <unordered_map>
#include <vector>

namespace mrt {

struct F3 {
    float x, y, z;
};

struct MeshData {
    std::vector<float> positions;  // xyzw per vertex (w = 0)
    std::vector<float> normals;    // xyzw per vertex
    std::vector<float> uvs;        // uv per vertex; empty if none
    std::vector<float> tangents;   // xyzw per vertex (w = handedness); empty if no uvs
    std::vector<uint32_t> indices;
    bool blasDirty = true;
    uint32_t vertexCount() const { return uint32_t(positions.size() / 4); }
    uint32_t triCount() const { return uint32_t(indices.size() / 3); }
};

struct InstanceData {
    uint32_t mesh = 0;
    uint32_t material = 0;
    float objectToWorld[16];  // column-major 4x4
};

struct MaterialData {
    float baseColor[3] = { 0.8f, 0.8f, 0.8f };
    float opacity = 1.0f;
    float roughness = 0.5f;
    float metallic = 0.0f;
    float transmission = 0.0f;
    float ior = 1.45f;
    float emission[3] = { 0.0f, 0.0f, 0.0f };
    uint32_t baseColorTex = 0, roughnessTex = 0, metallicTex = 0, normalTex = 0, emissionTex = 0;
};

struct TextureData {
    uint32_t width = 0, height = 0;
    uint32_t format = 0;            // mrtTextureFormat
    bool generateMips = false;
    std::vector<uint8_t> data;      // empty: removed slot
    uint32_t generation = 0;        // bumped on add/remove of this slot
};

struct LightData {
    uint32_t type = MRT_LIGHT_SUN;
    float direction[3] = { 0.0f, 1.0f, 0.0f };
    float position[3] = { 0.0f, 0.0f, 0.0f };
    float radiance[3] = { 1.0f, 1.0f, 1.0f };
    float angularRadius = 0.00465f;
    float corner[3] = { 0.0f, 0.0f, 0.0f };
    float edge0[3] = { 1.0f, 0.0f, 0.0f };
    float edge1[3] = { 0.0f, 0.0f, 1.0f };
    bool twoSided = false;
};

struct CameraData {
    float position[3] = { 0.0f, 0.0f, 5.0f };
    float target[3] = { 0.0f, 0.0f, 0.0f };
    float up[3] = { 0.0f, 1.0f, 0.0f };
    float fovYDeg = 45.0f;
};

struct CameraExData {
    int32_t projection = MRT_CAMERA_PERSPECTIVE;
    float position[3] = { 0.0f, 0.0f, 5.0f };
    float forward[3] = { 0.0f, 0.0f, -1.0f };
    float up[3] = { 0.0f, 1.0f, 0.0f };
    float left = -1.0f, right = 1.0f, bottom = -1.0f, top = 1.0f;
};

enum Dirty : uint32_t {
    kDirtyNone = 0,
    kDirtyGeometry = 1u << 0,
    kDirtyInstances = 1u << 1,
    kDirtyMaterials = 1u << 2,
    kDirtyLights = 1u << 3,
    kDirtyEnv = 1u << 4,
    kDirtyCamera = 1u << 5,
    kDirtyTextures = 1u << 6,
    kDirtyEnvParams = 1u << 7,
};

constexpr uint32_t kMaxTextureSlots = 64;

class Scene {
public:
    uint32_t addMesh(const mrtMeshDesc& d);
    void updateMesh(uint32_t id, const mrtMeshDesc& d);
    void removeMesh(uint32_t id);

    uint32_t addInstance(uint32_t mesh, const float* xform3x4, uint32_t material);
    void setInstanceTransform(uint32_t id, const float* xform3x4);
    void removeInstance(uint32_t id);

    uint32_t addMaterial(const mrtMaterialDesc& d);
    void updateMaterial(uint32_t id, const mrtMaterialDesc& d);
    void removeMaterial(uint32_t id);

    uint32_t addTexture(const mrtTextureDesc& d);
    void removeTexture(uint32_t id);
    uint32_t textureSlot(uint32_t id) const;  // ~0u when unknown

    uint32_t addLight(const mrtLightDesc& d);
    void updateLight(uint32_t id, const mrtLightDesc& d);
    void removeLight(uint32_t id);

    void setEnvironment(const mrtTextureDesc& hdri, float rotation, float intensity);
    void clearEnvironment();
    void setCamera(const mrtCameraDesc& d);
    void setCameraEx(const mrtCameraDescEx& d);

    uint32_t dirty() const { return dirty_; }
    void clearDirty() { dirty_ = kDirtyNone; }
    void markDirty(uint32_t bits) { dirty_ |= bits; }

    // Sorted views keep packing deterministic.
    std::map<uint32_t, MeshData>& meshes() { return meshes_; }
    const std::map<uint32_t, InstanceData>& instances() const { return instances_; }
    const std::map<uint32_t, MaterialData>& materials() const { return materials_; }
    const std::map<uint32_t, LightData>& lights() const { return lights_; }
    const std::vector<TextureData>& textures() const { return textures_; }
    const CameraData& camera() const { return camera_; }
    const CameraExData& cameraEx() const { return cameraEx_; }
    bool usesCameraEx() const { return useCameraEx_; }
    bool hasEnvironment() const { return envValid_; }
    const TextureData& environment() const { return env_; }
    float envRotation() const { return envRotation_; }
    float envIntensity() const { return envIntensity_; }

private:
    static void copyMesh(const mrtMeshDesc& d, MeshData& out);

    uint32_t nextId_ = 1;
    uint32_t dirty_ = kDirtyNone;
    std::map<uint32_t, MeshData> meshes_;
    std::map<uint32_t, InstanceData> instances_;
    std::map<uint32_t, MaterialData> materials_;
    std::map<uint32_t, LightData> lights_;
    std::vector<TextureData> textures_;                 // index = slot
    std::unordered_map<uint32_t, uint32_t> textureSlots_;  // id -> slot
    std::vector<uint32_t> freeSlots_;
    CameraData camera_;
    CameraExData cameraEx_;
    bool useCameraEx_ = false;
    TextureData env_;
    bool envValid_ = false;
    float envRotation_ = 0.0f;
    float envIntensity_ = 1.0f;
};

// Row-major 3x4 [R|t] (NULL = identity) -> column-major 4x4.
void mat4FromRows3x4(const float* rows, float out[16]);

} // namespace mrt
