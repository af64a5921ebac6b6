// Incremental scene container; see scene.h.
#include "scene.h"

#include <cmath>
#include <cstring>

namespace mrt {

namespace {

struct V {
    float x, y, z;
};
inline V sub(V a, V b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
inline V add(V a, V b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
inline V scale(V a, float s) { return { a.x * s, a.y * s, a.z * s }; }
inline V cross(V a, V b) {
    return { a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y };
}
inline float dot(V a, V b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline float len(V a) { return std::sqrt(dot(a, a)); }
inline V normalized(V a) { return scale(a, 1.0f / std::sqrt(dot(a, a))); }
inline V at4(const std::vector<float>& v, uint32_t i) { return { v[i * 4], v[i * 4 + 1], v[i * 4 + 2] }; }

// Area-weighted vertex normals (reference Scene.cpp:13-27).
void computeNormals(MeshData& m) {
    const uint32_t nv = m.vertexCount();
    std::vector<V> acc(nv, V{ 0.0f, 0.0f, 0.0f });
    for (size_t t = 0; t + 2 < m.indices.size(); t += 3) {
        const V p0 = at4(m.positions, m.indices[t]), p1 = at4(m.positions, m.indices[t + 1]),
                p2 = at4(m.positions, m.indices[t + 2]);
        const V n = cross(sub(p1, p0), sub(p2, p0));  // length = 2 * area
        for (int k = 0; k < 3; ++k) acc[m.indices[t + k]] = add(acc[m.indices[t + k]], n);
    }
    m.normals.assign(size_t(nv) * 4, 0.0f);
    for (uint32_t i = 0; i < nv; ++i) {
        const V n = len(acc[i]) > 1e-12f ? normalized(acc[i]) : V{ 0.0f, 1.0f, 0.0f };
        m.normals[i * 4] = n.x;
        m.normals[i * 4 + 1] = n.y;
        m.normals[i * 4 + 2] = n.z;
    }
}

// Per-vertex tangents: per-face accumulation then Gram-Schmidt against the
// normal, handedness in w (reference Scene.cpp:31-60).
void computeTangents(MeshData& m) {
    if (m.uvs.empty()) return;
    const uint32_t nv = m.vertexCount();
    std::vector<V> tan(nv, V{ 0.0f, 0.0f, 0.0f }), bit(nv, V{ 0.0f, 0.0f, 0.0f });
    for (size_t t = 0; t + 2 < m.indices.size(); t += 3) {
        const uint32_t i0 = m.indices[t], i1 = m.indices[t + 1], i2 = m.indices[t + 2];
        const V e1 = sub(at4(m.positions, i1), at4(m.positions, i0));
        const V e2 = sub(at4(m.positions, i2), at4(m.positions, i0));
        const float d1x = m.uvs[i1 * 2] - m.uvs[i0 * 2], d1y = m.uvs[i1 * 2 + 1] - m.uvs[i0 * 2 + 1];
        const float d2x = m.uvs[i2 * 2] - m.uvs[i0 * 2], d2y = m.uvs[i2 * 2 + 1] - m.uvs[i0 * 2 + 1];
        const float det = d1x * d2y - d2x * d1y;
        if (std::fabs(det) < 1e-12f) continue;
        const float r = 1.0f / det;
        const V T = scale(sub(scale(e1, d2y), scale(e2, d1y)), r);
        const V B = scale(sub(scale(e2, d1x), scale(e1, d2x)), r);
        for (uint32_t i : { i0, i1, i2 }) {
            tan[i] = add(tan[i], T);
            bit[i] = add(bit[i], B);
        }
    }
    m.tangents.assign(size_t(nv) * 4, 0.0f);
    for (uint32_t i = 0; i < nv; ++i) {
        const V n = at4(m.normals, i);
        V t = sub(tan[i], scale(n, dot(n, tan[i])));
        if (len(t) < 1e-9f)
            t = std::fabs(n.x) < 0.9f ? cross(n, V{ 1.0f, 0.0f, 0.0f }) : cross(n, V{ 0.0f, 1.0f, 0.0f });
        t = normalized(t);
        const float handed = dot(cross(n, t), bit[i]) < 0.0f ? -1.0f : 1.0f;
        m.tangents[i * 4] = t.x;
        m.tangents[i * 4 + 1] = t.y;
        m.tangents[i * 4 + 2] = t.z;
        m.tangents[i * 4 + 3] = handed;
    }
}

void identity4(float m[16]) {
    for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
}

bool meshDescValid(const mrtMeshDesc& d) {
    if (!d.positions || !d.indices || d.indexCount % 3 != 0) return false;
    for (uint32_t i = 0; i < d.indexCount; ++i)
        if (d.indices[i] >= d.vertexCount) return false;
    return true;
}

} // namespace

void mat4FromRows3x4(const float* r, float out[16]) {
    identity4(out);
    if (!r) return;
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 4; ++col) out[col * 4 + row] = r[row * 4 + col];
}

void Scene::copyMesh(const mrtMeshDesc& d, MeshData& out) {
    out.positions.assign(size_t(d.vertexCount) * 4, 0.0f);
    for (uint32_t i = 0; i < d.vertexCount; ++i)
        for (int k = 0; k < 3; ++k) out.positions[i * 4 + k] = d.positions[i * 3 + k];
    out.indices.assign(d.indices, d.indices + d.indexCount);
    if (d.normals) {
        out.normals.assign(size_t(d.vertexCount) * 4, 0.0f);
        for (uint32_t i = 0; i < d.vertexCount; ++i)
            for (int k = 0; k < 3; ++k) out.normals[i * 4 + k] = d.normals[i * 3 + k];
    } else {
        computeNormals(out);
    }
    if (d.uvs) {
        out.uvs.assign(d.uvs, d.uvs + size_t(d.vertexCount) * 2);
        computeTangents(out);
    } else {
        out.uvs.clear();
        out.tangents.clear();
    }
    out.blasDirty = true;
}

uint32_t Scene::addMesh(const mrtMeshDesc& d) {
    if (!meshDescValid(d)) return MRT_INVALID_ID;
    const uint32_t id = nextId_++;
    copyMesh(d, meshes_[id]);
    markDirty(kDirtyGeometry);
    return id;
}

void Scene::updateMesh(uint32_t id, const mrtMeshDesc& d) {
    auto it = meshes_.find(id);
    if (it == meshes_.end() || !meshDescValid(d)) return;
    copyMesh(d, it->second);
    markDirty(kDirtyGeometry);
}

void Scene::removeMesh(uint32_t id) {
    if (meshes_.erase(id) == 0) return;
    for (auto it = instances_.begin(); it != instances_.end();)
        it = (it->second.mesh == id) ? instances_.erase(it) : std::next(it);
    markDirty(kDirtyGeometry | kDirtyInstances);
}

uint32_t Scene::addInstance(uint32_t mesh, const float* xform, uint32_t material) {
    if (!meshes_.count(mesh)) return MRT_INVALID_ID;
    const uint32_t id = nextId_++;
    InstanceData& inst = instances_[id];
    inst.mesh = mesh;
    inst.material = material;
    mat4FromRows3x4(xform, inst.objectToWorld);
    markDirty(kDirtyInstances);
    return id;
}

void Scene::setInstanceTransform(uint32_t id, const float* xform) {
    auto it = instances_.find(id);
    if (it == instances_.end() || !xform) return;
    mat4FromRows3x4(xform, it->second.objectToWorld);
    markDirty(kDirtyInstances);
}

void Scene::removeInstance(uint32_t id) {
    if (instances_.erase(id)) markDirty(kDirtyInstances);
}

static MaterialData toMaterial(const mrtMaterialDesc& d) {
    MaterialData m;
    std::memcpy(m.baseColor, d.baseColor, sizeof(m.baseColor));
    m.opacity = d.opacity;
    m.roughness = d.roughness;
    m.metallic = d.metallic;
    m.transmission = d.transmission;
    m.ior = d.ior;
    std::memcpy(m.emission, d.emission, sizeof(m.emission));
    m.baseColorTex = d.baseColorTex;
    m.roughnessTex = d.roughnessTex;
    m.metallicTex = d.metallicTex;
    m.normalTex = d.normalTex;
    m.emissionTex = d.emissionTex;
    return m;
}

uint32_t Scene::addMaterial(const mrtMaterialDesc& d) {
    const uint32_t id = nextId_++;
    materials_[id] = toMaterial(d);
    markDirty(kDirtyMaterials);
    return id;
}

void Scene::updateMaterial(uint32_t id, const mrtMaterialDesc& d) {
    auto it = materials_.find(id);
    if (it == materials_.end()) return;
    it->second = toMaterial(d);
    markDirty(kDirtyMaterials);
}

void Scene::removeMaterial(uint32_t id) {
    if (materials_.erase(id)) markDirty(kDirtyMaterials | kDirtyInstances);
}

uint32_t Scene::addTexture(const mrtTextureDesc& d) {
    if (!d.pixels || d.width == 0 || d.height == 0 || uint32_t(d.format) > MRT_TEX_RGBA32F)
        return MRT_INVALID_ID;
    const uint32_t id = nextId_++;
    const size_t px = (d.format == MRT_TEX_RGBA32F) ? 16 : 4;
    TextureData t;
    t.width = d.width;
    t.height = d.height;
    t.format = uint32_t(d.format);
    t.generateMips = d.generateMips != 0;
    t.data.resize(size_t(d.width) * d.height * px);
    std::memcpy(t.data.data(), d.pixels, t.data.size());
    uint32_t slot;
    if (!freeSlots_.empty()) {
        slot = freeSlots_.back();
        freeSlots_.pop_back();
        t.generation = textures_[slot].generation + 1;
        textures_[slot] = std::move(t);
    } else {
        slot = uint32_t(textures_.size());
        t.generation = 1;
        textures_.push_back(std::move(t));
    }
    textureSlots_[id] = slot;
    markDirty(kDirtyTextures);
    return id;
}

void Scene::removeTexture(uint32_t id) {
    auto it = textureSlots_.find(id);
    if (it == textureSlots_.end()) return;
    const uint32_t slot = it->second;
    TextureData& t = textures_[slot];
    const uint32_t gen = t.generation + 1;
    t = TextureData{};
    t.generation = gen;
    textureSlots_.erase(it);
    freeSlots_.push_back(slot);
    markDirty(kDirtyTextures);
}

uint32_t Scene::textureSlot(uint32_t id) const {
    auto it = textureSlots_.find(id);
    return it != textureSlots_.end() ? it->second : ~0u;
}

static LightData toLight(const mrtLightDesc& d) {
    LightData l;
    l.type = uint32_t(d.type);
    std::memcpy(l.direction, d.direction, sizeof(l.direction));
    std::memcpy(l.position, d.position, sizeof(l.position));
    std::memcpy(l.radiance, d.radiance, sizeof(l.radiance));
    l.angularRadius = d.angularRadius;
    std::memcpy(l.corner, d.corner, sizeof(l.corner));
    std::memcpy(l.edge0, d.edge0, sizeof(l.edge0));
    std::memcpy(l.edge1, d.edge1, sizeof(l.edge1));
    l.twoSided = d.twoSided != 0;
    return l;
}

uint32_t Scene::addLight(const mrtLightDesc& d) {
    const uint32_t id = nextId_++;
    lights_[id] = toLight(d);
    markDirty(kDirtyLights);
    return id;
}

void Scene::updateLight(uint32_t id, const mrtLightDesc& d) {
    auto it = lights_.find(id);
    if (it == lights_.end()) return;
    it->second = toLight(d);
    markDirty(kDirtyLights);
}

void Scene::removeLight(uint32_t id) {
    if (lights_.erase(id)) markDirty(kDirtyLights);
}

void Scene::setEnvironment(const mrtTextureDesc& hdri, float rotation, float intensity) {
    if (!hdri.pixels || hdri.format != MRT_TEX_RGBA32F || hdri.width == 0 || hdri.height == 0) return;
    env_.width = hdri.width;
    env_.height = hdri.height;
    env_.format = MRT_TEX_RGBA32F;
    env_.data.resize(size_t(hdri.width) * hdri.height * 16);
    std::memcpy(env_.data.data(), hdri.pixels, env_.data.size());
    envValid_ = true;
    envRotation_ = rotation;
    envIntensity_ = intensity;
    markDirty(kDirtyEnv | kDirtyLights);
}

void Scene::clearEnvironment() {
    envValid_ = false;
    env_.data.clear();
    markDirty(kDirtyEnv | kDirtyLights);
}

void Scene::setCamera(const mrtCameraDesc& d) {
    std::memcpy(camera_.position, d.position, sizeof(camera_.position));
    std::memcpy(camera_.target, d.target, sizeof(camera_.target));
    std::memcpy(camera_.up, d.up, sizeof(camera_.up));
    camera_.fovYDeg = d.fovYDeg;
    useCameraEx_ = false;
    markDirty(kDirtyCamera);
}

void Scene::setCameraEx(const mrtCameraDescEx& d) {
    cameraEx_.projection = d.projection;
    std::memcpy(cameraEx_.position, d.position, sizeof(cameraEx_.position));
    std::memcpy(cameraEx_.forward, d.forward, sizeof(cameraEx_.forward));
    std::memcpy(cameraEx_.up, d.up, sizeof(cameraEx_.up));
    cameraEx_.left = d.left;
    cameraEx_.right = d.right;
    cameraEx_.bottom = d.bottom;
    cameraEx_.top = d.top;
    useCameraEx_ = true;
    markDirty(kDirtyCamera);
}

} // namespace mrt
