/*
 * crisoptix.hpp -- the C++ host API of CrisOptix (Renderer / Scene / Camera / Pipeline /
 * ShaderBindingTable) over the C ABI of crisoptix.h.  Header only; link libcrisoptix_amd.so.
 *
 * The classes keep the reference's names, constructor arguments and call order, so a CrisOptix
 * host moves over by swapping includes:
 *   Renderer(w, h, ctx, Camera&, const Scene&), LaunchFrame(out), GetFrameIdx()
 *       include/graph/Renderer.h:8-21
 *   Scene(ctx, model instances, explicit emitters), GetSceneSize()   include/model/Scene.h:12-24
 *   ModelInstance(model, transform), Emitter(mesh, transform, color)
 *       include/model/ModelInstance.h, include/model/Emitter.h:58-73, src/Emitter.cpp:3-62
 *   Camera(pos, viewDir, up, aspectRatio, fov), GetData(), RotateAroundOrigin()
 *       include/graph/Camera.h:25-79, src/Camera.cpp:20-38
 *   Pipeline::AddRaygenShader / CreatePipeline                        include/graph/Pipeline.h:116-156
 *   ShaderBindingTable::AddCallableEntry / CreateSBT / GetSBT      include/graph/ShaderBindingTable.h:22-69
 * Differences of substance: there is no OptiX context (a Context wraps cx_ctx, one device);
 * device programs are kernels selected by id, so a Pipeline holds the raygen passes it launches
 * and a ShaderBindingTable holds materials, whose dense ids the triangles carry in place of the
 * callable indices of MeshSbtData (include/model/Mesh.h:10-16); the scene is one flattened
 * instance (rehydration reads transforms[0], include/graph/Interaction.h:143).
 *
 * Errors: every failing cx_* call throws crisoptix::Error (a std::runtime_error with the
 * cx_status and cx_last_error's message), what the reference's ASSERT_CUDA / ASSERT_OPTIX do in
 * debug builds (include/util/custom_assert.h:26-66).
 */
#ifndef CRISOPTIX_HPP
#define CRISOPTIX_HPP

#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "crisoptix.h"

namespace crisoptix {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), m_Status(status) {}
    int Status() const { return m_Status; }

private:
    int m_Status;
};

inline void Check(int status, const cx_ctx* ctx, const char* call) {
    if (status == CX_OK) return;
    throw Error(status, std::string(call) + " failed (" + std::to_string(status) + "): " +
                            (ctx ? cx_last_error(ctx) : "no context"));
}

// glm::vec3 of the reference's interfaces.
struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};

// InitCuda + InitOptix (src/main.cpp:35-65): one device.
class Context {
public:
    explicit Context(int device = 0) {
        const int rc = cx_ctx_create(device, &m_Ctx);
        if (rc != CX_OK) throw Error(rc, "cx_ctx_create failed: no usable HIP device " + std::to_string(device));
        if (cx_abi_version() != CX_ABI_VERSION)
            throw Error(CX_ERR_UNSUPPORTED, "libcrisoptix_amd.so was built for another crisoptix.h");
    }
    ~Context() { cx_ctx_destroy(m_Ctx); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    cx_ctx* Get() const { return m_Ctx; }

private:
    cx_ctx* m_Ctx = nullptr;
};

// Material constants and texture bindings (include/model/Material.h:16-33).
struct Material {
    Vec3 diffColor{0.5f, 0.5f, 0.5f};
    Vec3 specF0{0.0f, 0.0f, 0.0f};
    Vec3 emissiveColor{0.0f, 0.0f, 0.0f};
    float roughness = 1.0f;
    int32_t diffTex = -1, specTex = -1, roughTex = -1;   // indices into Scene's textures
    cx_spec_tex_usage specTexUsage = CX_SPEC_TEX_COLOR;
};

// The reference registers a material's eval and sample callables and stores their SBT indices
// in each mesh's hit record (src/Material.cpp:36-43, src/Mesh.cpp:27-32, 137-146).  Here the
// BSDF is one inline function of the kernels, so an entry is the material itself and its index
// is the dense material id the triangles carry.
class ShaderBindingTable {
public:
    // Adds a material; returns its id (the reference's callable index).
    uint32_t AddCallableEntry(const Material& m) {
        if (m_Created) throw Error(CX_ERR_INVALID_ARG, "ShaderBindingTable: entry added after CreateSBT");
        cx_material row{};
        const Vec3* src[3] = {&m.diffColor, &m.specF0, &m.emissiveColor};
        float* dst[3] = {row.diffColor, row.specF0, row.emissiveColor};
        for (int i = 0; i < 3; i++) {
            dst[i][0] = src[i]->x;
            dst[i][1] = src[i]->y;
            dst[i][2] = src[i]->z;
        }
        row.roughness = m.roughness;
        m_Rows.push_back(row);
        m_Tex.push_back(cx_material_tex{m.diffTex, m.specTex, m.roughTex, (uint32_t)m.specTexUsage});
        return (uint32_t)(m_Rows.size() - 1);
    }
    void CreateSBT() {
        if (m_Rows.empty()) throw Error(CX_ERR_INVALID_ARG, "ShaderBindingTable: no material");
        m_Created = true;
    }
    bool Created() const { return m_Created; }
    const std::vector<cx_material>& Rows() const { return m_Rows; }
    const std::vector<cx_material_tex>& Textures() const { return m_Tex; }

private:
    std::vector<cx_material> m_Rows;
    std::vector<cx_material_tex> m_Tex;
    bool m_Created = false;
};

// A mesh of the flattened instance (src/Mesh.cpp: Vertex pos/normal/uv, include/model/Vertex.h;
// 3 indices per triangle) and the id of its material.  `visibility` is the mesh's CX_VIS_* mask
// (the OptiX instance visibilityMask): a light's own geometry is CX_VIS_SURFACE.
struct Mesh {
    std::vector<float> positions;   // [vertices][3]
    std::vector<float> normals;     // [vertices][3]
    std::vector<float> uvs;         // [vertices][2], or empty (all zero)
    std::vector<uint32_t> indices;  // [triangles][3]
    uint32_t materialId = 0;
    uint8_t visibility = CX_VIS_ALL;
};

// glm::mat4: column-major, m[4 * column + row].
struct Mat4 {
    float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    // (M * vec4(p, 1)).xyz, summed column by column in float as glm's operator* does.
    Vec3 TransformPoint(const float* p) const {
        float o[3];
        for (int r = 0; r < 3; r++) o[r] = ((m[r] * p[0] + m[4 + r] * p[1]) + m[8 + r] * p[2]) + m[12 + r];
        return Vec3{o[0], o[1], o[2]};
    }
    // mat3(M) * d
    Vec3 TransformDir(const float* d) const {
        float o[3];
        for (int r = 0; r < 3; r++) o[r] = (m[r] * d[0] + m[4 + r] * d[1]) + m[8 + r] * d[2];
        return Vec3{o[0], o[1], o[2]};
    }
};

// Model (include/model/Model.h): its meshes, in the model's own space.  Loading from files is the
// caller's (crisoptix_amd/assets.py reads OBJ/MTL as the reference's assimp path does).
struct Model {
    std::vector<Mesh> meshes;
};

// ModelInstance(model, transform) (include/model/ModelInstance.h): a model placed in the world.
class ModelInstance {
public:
    ModelInstance(const Model& model, const Mat4& transform) : m_Model(model), m_Transform(transform) {}
    const Model& GetModel() const { return m_Model; }
    const Mat4& GetTransform() const { return m_Transform; }
    void SetTransform(const Mat4& transform) { m_Transform = transform; }

private:
    const Model& m_Model;
    Mat4 m_Transform;
};

// Emitter (include/model/Emitter.h:58-73, src/Emitter.cpp:3-62): world-space faces and one
// colour; built from a mesh and its transform as the reference's constructor is (the vertices
// transformed on the host), or from faces already in world space.
struct Emitter {
    std::vector<float> faces;       // [faces][3 vertices][3]
    Vec3 color{1.0f, 1.0f, 1.0f};

    Emitter() = default;
    Emitter(std::vector<float> worldFaces, const Vec3& c) : faces(std::move(worldFaces)), color(c) {}
    Emitter(const Mesh* mesh, const Mat4& transform, const Vec3& c) : color(c) {
        if (!mesh || mesh->indices.size() % 3 != 0) throw Error(CX_ERR_INVALID_ARG, "Emitter: mesh without triangles");
        const size_t nv = mesh->positions.size() / 3;
        for (uint32_t i : mesh->indices) {
            if (i >= nv) throw Error(CX_ERR_INVALID_ARG, "Emitter: vertex index out of range");
            const Vec3 w = transform.TransformPoint(&mesh->positions[3 * (size_t)i]);
            faces.insert(faces.end(), {w.x, w.y, w.z});
        }
    }
};

// An RGBA8 texture (src/Texture.cpp:7-76).
struct Texture {
    std::vector<uint8_t> rgba;      // [height][width][4]
    uint32_t width = 0, height = 0;
};

struct SceneOptions {
    uint32_t bvhWidth = 0;          // 0: the library's default (4)
    uint32_t bvhMaxLeafSize = 0;    // 0: the library's default (2)
    cx_bvh_builder bvhBuilder = CX_BVH_BUILD_HOST;
};

// Scene(ctx, instances, emitters) (include/model/Scene.h:12-24; src/Scene.cpp:4-204): flattens
// the meshes into one triangle list with per-triangle material ids and builds the BVH.
class Scene {
public:
    // Scene(ctx, modelInstances, explicitEmitters) (include/model/Scene.h:12-24): every instance's
    // meshes flattened into one instance with the transforms applied on the host (positions by
    // M, normals by mat3(M), as rehydration reads one transform, include/graph/Interaction.h:143).
    Scene(const Context& ctx, const std::vector<const ModelInstance*>& instances, const ShaderBindingTable& sbt,
          const std::vector<const Emitter*>& emitters, const std::vector<Texture>& textures = {},
          const SceneOptions& opt = {}) {
        std::vector<Mesh> world;
        for (const ModelInstance* inst : instances) {
            if (!inst) throw Error(CX_ERR_INVALID_ARG, "Scene: null model instance");
            const Mat4& M = inst->GetTransform();
            for (const Mesh& m : inst->GetModel().meshes) {
                Mesh w = m;
                for (size_t v = 0; 3 * v + 2 < m.positions.size(); v++) {
                    const Vec3 p = M.TransformPoint(&m.positions[3 * v]);
                    w.positions[3 * v] = p.x;
                    w.positions[3 * v + 1] = p.y;
                    w.positions[3 * v + 2] = p.z;
                }
                for (size_t v = 0; 3 * v + 2 < m.normals.size(); v++) {
                    const Vec3 n = M.TransformDir(&m.normals[3 * v]);
                    w.normals[3 * v] = n.x;
                    w.normals[3 * v + 1] = n.y;
                    w.normals[3 * v + 2] = n.z;
                }
                world.push_back(std::move(w));
            }
        }
        std::vector<Emitter> em;
        for (const Emitter* e : emitters) {
            if (!e) throw Error(CX_ERR_INVALID_ARG, "Scene: null emitter");
            em.push_back(*e);
        }
        Build(ctx, world, sbt, em, textures, opt);
    }

    // The same from meshes and emitters already in world space.
    Scene(const Context& ctx, const std::vector<Mesh>& meshes, const ShaderBindingTable& sbt,
          const std::vector<Emitter>& emitters, const std::vector<Texture>& textures = {},
          const SceneOptions& opt = {}) {
        Build(ctx, meshes, sbt, emitters, textures, opt);
    }

    ~Scene() { cx_scene_destroy(m_Scene); }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    // Scene::GetSceneSize (src/Scene.cpp:48-71): diagonal of the vertices' box and the origin.
    float GetSceneSize() const { return GetInfo().sceneSize; }
    cx_scene_info GetInfo() const {
        cx_scene_info info{};
        Check(cx_scene_get_info(m_Scene, &info), m_Ctx, "cx_scene_get_info");
        return info;
    }
    cx_scene* Get() const { return m_Scene; }

private:
    void Build(const Context& ctx, const std::vector<Mesh>& meshes, const ShaderBindingTable& sbt,
               const std::vector<Emitter>& emitters, const std::vector<Texture>& textures, const SceneOptions& opt) {
        if (!sbt.Created()) throw Error(CX_ERR_INVALID_ARG, "Scene: ShaderBindingTable::CreateSBT not called");
        std::vector<float> pos, nrm, uv;
        std::vector<uint32_t> idx, triMat;
        std::vector<uint8_t> vis;
        bool anyUv = false, anyMask = false;
        for (const Mesh& m : meshes) anyUv |= !m.uvs.empty();
        for (const Mesh& m : meshes) {
            const size_t nv = m.positions.size() / 3;
            if (m.positions.size() != 3 * nv || m.normals.size() != 3 * nv || m.indices.size() % 3 != 0 ||
                (!m.uvs.empty() && m.uvs.size() != 2 * nv))
                throw Error(CX_ERR_INVALID_ARG, "Scene: mesh arrays of inconsistent sizes");
            if (m.materialId >= sbt.Rows().size()) throw Error(CX_ERR_INVALID_ARG, "Scene: unknown material id");
            const uint32_t base = (uint32_t)(pos.size() / 3);
            pos.insert(pos.end(), m.positions.begin(), m.positions.end());
            nrm.insert(nrm.end(), m.normals.begin(), m.normals.end());
            if (anyUv) {
                if (m.uvs.empty()) uv.insert(uv.end(), 2 * nv, 0.0f);
                else uv.insert(uv.end(), m.uvs.begin(), m.uvs.end());
            }
            for (uint32_t i : m.indices) idx.push_back(base + i);
            triMat.insert(triMat.end(), m.indices.size() / 3, m.materialId);
            vis.insert(vis.end(), m.indices.size() / 3, m.visibility);
            anyMask |= m.visibility != CX_VIS_ALL;
        }
        std::vector<float> faces, colors;
        std::vector<uint32_t> offsets{0};
        for (const Emitter& e : emitters) {
            if (e.faces.empty() || e.faces.size() % 9 != 0) throw Error(CX_ERR_INVALID_ARG, "Scene: emitter faces");
            faces.insert(faces.end(), e.faces.begin(), e.faces.end());
            offsets.push_back((uint32_t)(faces.size() / 9));
            colors.insert(colors.end(), {e.color.x, e.color.y, e.color.z});
        }
        std::vector<cx_texture> tex;
        for (const Texture& t : textures) tex.push_back(cx_texture{t.rgba.data(), t.width, t.height});
        cx_scene_desc d{};
        d.positions = pos.data();
        d.normals = nrm.data();
        d.numVertices = (uint32_t)(pos.size() / 3);
        d.indices = idx.data();
        d.triMaterial = triMat.data();
        d.numTriangles = (uint32_t)triMat.size();
        d.materials = sbt.Rows().data();
        d.numMaterials = (uint32_t)sbt.Rows().size();
        d.emitterFaces = faces.data();
        d.emitterFaceOffset = offsets.data();
        d.emitterColor = colors.data();
        d.numEmitters = (uint32_t)emitters.size();
        d.bvhMaxLeafSize = opt.bvhMaxLeafSize;
        d.bvhWidth = opt.bvhWidth;
        d.uvs = anyUv ? uv.data() : nullptr;
        d.materialTex = sbt.Textures().data();
        d.textures = tex.empty() ? nullptr : tex.data();
        d.numTextures = (uint32_t)tex.size();
        d.bvhBuilder = (uint32_t)opt.bvhBuilder;
        d.triMask = anyMask ? vis.data() : nullptr;
        Check(cx_scene_build(ctx.Get(), &d, &m_Scene), ctx.Get(), "cx_scene_build");
        m_Ctx = ctx.Get();
    }

    cx_scene* m_Scene = nullptr;
    cx_ctx* m_Ctx = nullptr;
};

// Camera (include/graph/Camera.h:25-79, src/Camera.cpp): position, view direction, up vector,
// aspect ratio and vertical field of view (radians).  GetData() gives the pinhole basis of
// SpawnCameraRay (include/graph/Camera.h:14-22): W = the normalised view direction, U = s *
// tan(fov / 2) * aspect and V = -u * tan(fov / 2) for s = normalize(W x up), u = s x W (the rows of
// glm::lookAt), computed in double and rounded once.
class Camera {
public:
    Camera(const Vec3& pos, const Vec3& viewDir, const Vec3& up, float aspectRatio, float fov)
        : m_Pos(pos), m_ViewDir(viewDir), m_Up(up), m_AspectRatio(aspectRatio), m_Fov(fov) {}

    cx_camera GetData() const {
        const std::array<double, 3> f0{m_ViewDir.x, m_ViewDir.y, m_ViewDir.z}, up{m_Up.x, m_Up.y, m_Up.z};
        const std::array<double, 3> f = Normalize(f0);
        const std::array<double, 3> s = Normalize(Cross(f, up));
        const std::array<double, 3> u = Cross(s, f);
        const double un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        const double th = std::tan(0.5 * (double)m_Fov);
        cx_camera c{};
        c.pos[0] = m_Pos.x;
        c.pos[1] = m_Pos.y;
        c.pos[2] = m_Pos.z;
        for (int i = 0; i < 3; i++) {
            c.U[i] = (float)(s[i] * th * (double)m_AspectRatio);
            c.V[i] = (float)(-u[i] / un * th);
            c.W[i] = (float)f[i];
        }
        return c;
    }

    // Camera::RotateAroundOrigin (src/Camera.cpp:34-38): rotates the position about `axis`
    // through the origin and looks at the origin.
    void RotateAroundOrigin(const Vec3& axis, float angle) {
        const std::array<double, 3> a = Normalize({axis.x, axis.y, axis.z});
        const double c = std::cos((double)angle), s = std::sin((double)angle), x = a[0], y = a[1], z = a[2];
        const double R[3][3] = {{c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s},
                                {y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s},
                                {z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)}};
        const double p[3] = {m_Pos.x, m_Pos.y, m_Pos.z};
        float np[3];
        for (int i = 0; i < 3; i++) np[i] = (float)(R[i][0] * p[0] + R[i][1] * p[1] + R[i][2] * p[2]);
        m_Pos = Vec3{np[0], np[1], np[2]};
        const std::array<double, 3> d = Normalize({-(double)np[0], -(double)np[1], -(double)np[2]});
        m_ViewDir = Vec3{(float)d[0], (float)d[1], (float)d[2]};
        m_Changed = true;
    }

    const Vec3& GetPos() const { return m_Pos; }
    void SetPos(const Vec3& p) { m_Pos = p; m_Changed = true; }
    const Vec3& GetViewDir() const { return m_ViewDir; }
    void SetViewDir(const Vec3& d) { m_ViewDir = d; m_Changed = true; }
    const Vec3& GetUp() const { return m_Up; }
    void SetUp(const Vec3& u) { m_Up = u; m_Changed = true; }
    float GetAspectRatio() const { return m_AspectRatio; }
    void SetAspectRatio(float a) { m_AspectRatio = a; m_Changed = true; }
    void SetAspectRatio(uint32_t width, uint32_t height) { SetAspectRatio((float)width / (float)height); }
    float GetFov() const { return m_Fov; }
    void SetFov(float fov) { m_Fov = fov; m_Changed = true; }
    bool HasChanged() const { return m_Changed; }
    void SetChanged(bool changed) { m_Changed = changed; }

private:
    static std::array<double, 3> Cross(const std::array<double, 3>& a, const std::array<double, 3>& b) {
        return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    }
    static std::array<double, 3> Normalize(const std::array<double, 3>& a) {
        const double n = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        return {a[0] / n, a[1] / n, a[2] / n};
    }

    Vec3 m_Pos, m_ViewDir, m_Up;
    float m_AspectRatio = 1.0f;
    float m_Fov = 1.0471975511965976f;   // 60 degrees, Camera.h:71
    bool m_Changed = false;
};

class Renderer;

// A pipeline groups the raygen programs a frame launches (src/Renderer.cpp:76-153 builds one per
// raygen, each launched with optixLaunch at :259-371).  Here a raygen is a pass id: AddRaygenShader
// returns the pass's index in the pipeline, CreatePipeline freezes the list, and Launch runs one
// registered pass on a renderer's current frame (cx_pass_launch).
class Pipeline {
public:
    uint32_t AddRaygenShader(cx_pass pass) {
        if (m_Created) throw Error(CX_ERR_INVALID_ARG, "Pipeline: raygen added after CreatePipeline");
        if ((int)pass < 0 || pass >= CX_PASS_COUNT) throw Error(CX_ERR_INVALID_ARG, "Pipeline: unknown pass");
        m_Raygens.push_back(pass);
        return (uint32_t)(m_Raygens.size() - 1);
    }
    void CreatePipeline() {
        if (m_Raygens.empty()) throw Error(CX_ERR_INVALID_ARG, "Pipeline: no raygen");
        m_Created = true;
    }
    const std::vector<cx_pass>& Raygens() const { return m_Raygens; }
    inline void Launch(Renderer& renderer, uint32_t raygenIdx, const cx_launch_params& params,
                       void* stream = nullptr) const;

private:
    std::vector<cx_pass> m_Raygens;
    bool m_Created = false;
};

// Renderer(w, h, ctx, Camera&, const Scene&) and LaunchFrame (include/graph/Renderer.h:8-21,
// src/Renderer.cpp:13-406).  The parameter block starts at the reference's defaults
// (src/Renderer.cpp:34-58); GetLaunchParams / SetLaunchParams stand in for the ImGui panel that
// edits it (src/Renderer.cpp:413-470).
class Renderer {
public:
    Renderer(uint32_t width, uint32_t height, const Context& ctx, Camera& cam, const Scene& scene)
        : m_Width(width), m_Height(height), m_Ctx(ctx.Get()), m_Cam(cam), m_Scene(scene) {
        Check(cx_renderer_create(m_Ctx, scene.Get(), width, height, &m_Renderer), m_Ctx, "cx_renderer_create");
        cx_launch_params p{};
        Check(cx_default_params(scene.Get(), width, height, &p), m_Ctx, "cx_default_params");
        Check(cx_renderer_set_params(m_Renderer, &p), m_Ctx, "cx_renderer_set_params");
    }
    ~Renderer() { cx_renderer_destroy(m_Renderer); }
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    // One frame into outputBuffer, device memory of [height][width] float3 (the reference's
    // glm::vec3*); the camera's changed flag resets temporal history as at src/Renderer.cpp:213.
    void LaunchFrame(float* outputBuffer, void* stream = nullptr) {
        const cx_camera c = m_Cam.GetData();
        Check(cx_renderer_frame(m_Renderer, &c, m_Cam.HasChanged() ? 1 : 0, outputBuffer, stream), m_Ctx,
              "cx_renderer_frame");
        m_Cam.SetChanged(false);
    }

    // The frame fields of a frame whose passes the caller launches through a Pipeline.
    void BeginFrame(float* outputBuffer) {
        const cx_camera c = m_Cam.GetData();
        Check(cx_renderer_begin_frame(m_Renderer, &c, m_Cam.HasChanged() ? 1 : 0, outputBuffer), m_Ctx,
              "cx_renderer_begin_frame");
        m_Cam.SetChanged(false);
    }

    // Renderer::GetFrameIdx: frames accumulated since the last camera change (0 when
    // accumulation is off, src/Renderer.cpp:212-213).
    size_t GetFrameIdx() const { return (size_t)GetLaunchParams().frameIdx; }

    cx_launch_params GetLaunchParams() const {
        cx_launch_params p{};
        Check(cx_renderer_get_params(m_Renderer, &p), m_Ctx, "cx_renderer_get_params");
        return p;
    }
    void SetLaunchParams(const cx_launch_params& p) {
        Check(cx_renderer_set_params(m_Renderer, &p), m_Ctx, "cx_renderer_set_params");
    }
    void Seed(uint64_t seed) { Check(cx_renderer_seed(m_Renderer, seed), m_Ctx, "cx_renderer_seed"); }
    cx_frame_stats GetStats() {
        cx_frame_stats st{};
        Check(cx_renderer_get_stats(m_Renderer, &st), m_Ctx, "cx_renderer_get_stats");
        return st;
    }
    uint32_t GetWidth() const { return m_Width; }
    uint32_t GetHeight() const { return m_Height; }
    cx_renderer* Get() const { return m_Renderer; }
    cx_ctx* GetContext() const { return m_Ctx; }

private:
    uint32_t m_Width, m_Height;
    cx_ctx* m_Ctx;
    Camera& m_Cam;
    const Scene& m_Scene;
    cx_renderer* m_Renderer = nullptr;
};

inline void Pipeline::Launch(Renderer& renderer, uint32_t raygenIdx, const cx_launch_params& params,
                             void* stream) const {
    if (!m_Created) throw Error(CX_ERR_INVALID_ARG, "Pipeline: CreatePipeline not called");
    if (raygenIdx >= m_Raygens.size()) throw Error(CX_ERR_INVALID_ARG, "Pipeline: raygen index out of range");
    Check(cx_pass_launch(renderer.Get(), m_Raygens[raygenIdx], &params, stream), renderer.GetContext(),
          "cx_pass_launch");
}

}  // namespace crisoptix

#endif  // CRISOPTIX_HPP
