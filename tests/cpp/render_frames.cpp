// render_frames: renders a scene through the C++ host API (include/crisoptix.hpp) the way the
// reference's main loop drives Renderer::LaunchFrame (src/main.cpp:190-260), and once more
// through a Pipeline of raygen passes in the order of src/Renderer.cpp:259-371, writing every
// frame of both runs for the test that compares them with the oracle.
//
//   render_frames --abi                      prints "abi <header> <library>" (no GPU needed)
//   render_frames SCENE OUT W H FRAMES SEED [orbit=0|1] [instanced=0|1] [field=value ...]
//
// SCENE (little-endian): "CXS1", u32 numVertices numTriangles numMaterials numEmitters
//   numEmitterFaces numMeshes hasUv; f32 positions[V][3] normals[V][3] (uvs[V][2]);
//   u32 indices[T][3]; f32 materials[M][10]; f32 faces[F][9]; u32 faceOffset[E+1];
//   f32 emitterColor[E][3]; per mesh u32 firstVertex numVertices firstTriangle numTriangles
//   material visibility; camera f32 pos[3] viewDir[3] up[3] fov.
// OUT: f32 sceneSize, u32 frames; per frame cx_camera, u32 frameIdx, f32 image[H][W][3] of
//   LaunchFrame, f32 image[H][W][3] of the Pipeline run.
// field=value overrides a cx_launch_params field of the Renderer defaults (src/Renderer.cpp:34-58).
// instanced=1: the file's vertices and emitter faces are in a model space turned by -90 degrees
// about y; the scene is then built from one Model placed by a ModelInstance whose transform turns
// it back (exactly: the matrix holds only 0 and +-1), and each emitter from a mesh of its faces
// with the same transform, as Scene(ctx, instances, emitters) and Emitter(mesh, transform, colour)
// are used in the reference.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "crisoptix.hpp"

namespace {

struct Reader {
    std::ifstream in;
    explicit Reader(const char* path) : in(path, std::ios::binary) {
        if (!in) throw std::runtime_error(std::string("cannot open ") + path);
    }
    template <typename T>
    std::vector<T> read(size_t n) {
        std::vector<T> v(n);
        if (n && !in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T))))
            throw std::runtime_error("scene file truncated");
        return v;
    }
    uint32_t u32() { return read<uint32_t>(1)[0]; }
};

void set_field(cx_launch_params& p, const std::string& kv) {
    const size_t eq = kv.find('=');
    if (eq == std::string::npos) throw std::runtime_error("expected field=value: " + kv);
    const std::string k = kv.substr(0, eq);
    const double v = std::atof(kv.c_str() + eq + 1);
    struct F { const char* name; int32_t* i; float* f; };
    const F fields[] = {
        {"windowRadius", reinterpret_cast<int32_t*>(&p.windowRadius), nullptr},
        {"neeProb", nullptr, &p.neeProb},
        {"enableAccum", &p.enableAccum, nullptr},
        {"maxPathLen", &p.maxPathLen, nullptr},
        {"rendererType", &p.rendererType, nullptr},
        {"reconMinRoughness", nullptr, &p.reconMinRoughness},
        {"reconMinDistance", nullptr, &p.reconMinDistance},
        {"prefixLen", &p.prefixLen, nullptr},
        {"prefixEnableTemporal", &p.prefixEnableTemporal, nullptr},
        {"prefixEnableSpatial", &p.prefixEnableSpatial, nullptr},
        {"prefixSpatialCount", &p.prefixSpatialCount, nullptr},
        {"prefixSpatialRounds", &p.prefixSpatialRounds, nullptr},
        {"suffixEnableTemporal", &p.suffixEnableTemporal, nullptr},
        {"suffixEnableSpatial", &p.suffixEnableSpatial, nullptr},
        {"suffixSpatialCount", &p.suffixSpatialCount, nullptr},
        {"suffixSpatialRounds", &p.suffixSpatialRounds, nullptr},
        {"gatherN", &p.gatherN, nullptr},
        {"gatherM", &p.gatherM, nullptr},
        {"gatherRadius", nullptr, &p.gatherRadius},
        {"gatherRadiusType", &p.gatherRadiusType, nullptr},
    };
    for (const F& f : fields) {
        if (k != f.name) continue;
        if (f.i) *f.i = (int32_t)v;
        else *f.f = (float)v;
        return;
    }
    throw std::runtime_error("unknown launch parameter " + k);
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// The frame loop of Renderer::LaunchFrame (src/Renderer.cpp:259-371) as a caller would write it
// over a Pipeline whose raygens are the frame's passes.
void pipeline_frame(crisoptix::Renderer& ren, const crisoptix::Pipeline& pipe, const uint32_t rg[CX_PASS_COUNT],
                    float* out) {
    ren.BeginFrame(out);
    cx_launch_params p = ren.GetLaunchParams();
    const bool restir = p.rendererType == CX_RESTIR_PT || p.rendererType == CX_CONDITIONAL_RESTIR;
    const bool cond = p.rendererType == CX_CONDITIONAL_RESTIR;
    pipe.Launch(ren, rg[CX_PASS_PREFIX_GEN_TEMP], p);
    if (restir && p.prefixEnableSpatial) {
        for (int i = 0; i < p.prefixSpatialRounds; i++) {
            p.spatialRoundIdx = i;
            pipe.Launch(ren, rg[CX_PASS_PREFIX_SPATIAL], p);
        }
    }
    if (!cond) return;
    pipe.Launch(ren, rg[CX_PASS_SUFFIX_GEN_TEMP], p);
    if (p.suffixEnableSpatial && p.suffixSpatialCount > 0) {
        for (int i = 0; i < p.suffixSpatialRounds; i++) {
            p.spatialRoundIdx = i;
            pipe.Launch(ren, rg[CX_PASS_SUFFIX_SPATIAL], p);
        }
    }
    pipe.Launch(ren, rg[CX_PASS_PREFIX_STORE_ENTRIES], p);
    pipe.Launch(ren, rg[CX_PASS_ENTRIES_BUILD], p);
    pipe.Launch(ren, rg[CX_PASS_FINAL_GATHER], p);
}

int run(int argc, char** argv) {
    if (argc == 2 && std::string(argv[1]) == "--abi") {
        std::printf("abi %d %d\n", CX_ABI_VERSION, cx_abi_version());
        return 0;
    }
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s SCENE OUT W H FRAMES SEED [orbit=0|1] [field=value ...]\n", argv[0]);
        return 2;
    }
    const uint32_t W = (uint32_t)std::atoi(argv[3]), H = (uint32_t)std::atoi(argv[4]);
    const int frames = std::atoi(argv[5]);
    const uint64_t seed = std::strtoull(argv[6], nullptr, 10);
    bool orbit = false, instanced = false;
    std::vector<std::string> overrides;
    for (int i = 7; i < argc; i++) {
        const std::string a = argv[i];
        if (a.rfind("orbit=", 0) == 0) orbit = a == "orbit=1";
        else if (a.rfind("instanced=", 0) == 0) instanced = a == "instanced=1";
        else overrides.push_back(a);
    }

    Reader rd(argv[1]);
    const std::vector<char> magic = rd.read<char>(4);
    if (std::memcmp(magic.data(), "CXS1", 4) != 0) throw std::runtime_error("not a CXS1 scene file");
    const uint32_t nv = rd.u32(), nt = rd.u32(), nm = rd.u32(), ne = rd.u32(), nf = rd.u32(), nmesh = rd.u32(),
                   hasUv = rd.u32();
    const std::vector<float> pos = rd.read<float>(3 * (size_t)nv), nrm = rd.read<float>(3 * (size_t)nv);
    const std::vector<float> uv = rd.read<float>(hasUv ? 2 * (size_t)nv : 0);
    const std::vector<uint32_t> idx = rd.read<uint32_t>(3 * (size_t)nt);
    const std::vector<float> mats = rd.read<float>(10 * (size_t)nm), faces = rd.read<float>(9 * (size_t)nf);
    const std::vector<uint32_t> faceOff = rd.read<uint32_t>(ne + 1);
    const std::vector<float> ecol = rd.read<float>(3 * (size_t)ne);
    const std::vector<uint32_t> meshRec = rd.read<uint32_t>(6 * (size_t)nmesh);
    const std::vector<float> camRec = rd.read<float>(10);

    crisoptix::Context ctx(0);

    // Materials: one callable entry each (src/Material.cpp:36-43).
    crisoptix::ShaderBindingTable sbt;
    for (uint32_t m = 0; m < nm; m++) {
        const float* r = &mats[10 * (size_t)m];
        crisoptix::Material mat;
        mat.diffColor = {r[0], r[1], r[2]};
        mat.specF0 = {r[3], r[4], r[5]};
        mat.emissiveColor = {r[6], r[7], r[8]};
        mat.roughness = r[9];
        if (sbt.AddCallableEntry(mat) != m) throw std::runtime_error("material ids are not dense");
    }
    sbt.CreateSBT();

    // Meshes over their own vertex ranges, indices local to the mesh (src/Mesh.cpp:54-135).
    std::vector<crisoptix::Mesh> meshes(nmesh);
    for (uint32_t i = 0; i < nmesh; i++) {
        const uint32_t* m = &meshRec[6 * (size_t)i];
        const uint32_t v0 = m[0], vn = m[1], t0 = m[2], tn = m[3];
        if ((size_t)v0 + vn > nv || (size_t)t0 + tn > nt) throw std::runtime_error("mesh range outside the arrays");
        crisoptix::Mesh& me = meshes[i];
        me.positions.assign(pos.begin() + 3 * (size_t)v0, pos.begin() + 3 * (size_t)(v0 + vn));
        me.normals.assign(nrm.begin() + 3 * (size_t)v0, nrm.begin() + 3 * (size_t)(v0 + vn));
        if (hasUv) me.uvs.assign(uv.begin() + 2 * (size_t)v0, uv.begin() + 2 * (size_t)(v0 + vn));
        for (size_t k = 3 * (size_t)t0; k < 3 * (size_t)(t0 + tn); k++) {
            if (idx[k] < v0 || idx[k] >= v0 + vn) throw std::runtime_error("triangle outside its mesh");
            me.indices.push_back(idx[k] - v0);
        }
        me.materialId = m[4];
        me.visibility = (uint8_t)m[5];
    }
    std::vector<crisoptix::Emitter> emitters(ne);
    crisoptix::Mat4 turn;   // +90 degrees about y: (x, y, z) -> (z, y, -x)
    const float cols[16] = {0, 0, -1, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 1};
    std::memcpy(turn.m, cols, sizeof(cols));
    std::vector<crisoptix::Mesh> emitterMeshes(ne);
    for (uint32_t e = 0; e < ne; e++) {
        const crisoptix::Vec3 color{ecol[3 * e], ecol[3 * e + 1], ecol[3 * e + 2]};
        std::vector<float> f(faces.begin() + 9 * (size_t)faceOff[e], faces.begin() + 9 * (size_t)faceOff[e + 1]);
        if (instanced) {
            crisoptix::Mesh& em = emitterMeshes[e];
            em.positions = f;
            for (uint32_t v = 0; v < f.size() / 3; v++) em.indices.push_back(v);
            emitters[e] = crisoptix::Emitter(&em, turn, color);
        } else {
            emitters[e] = crisoptix::Emitter(std::move(f), color);
        }
    }
    crisoptix::Model model;
    model.meshes = meshes;
    const crisoptix::ModelInstance instance(model, turn);
    std::vector<const crisoptix::Emitter*> emitterPtrs;
    for (const crisoptix::Emitter& e : emitters) emitterPtrs.push_back(&e);
    crisoptix::Scene scene = instanced ? crisoptix::Scene(ctx, std::vector<const crisoptix::ModelInstance*>{&instance},
                                                          sbt, emitterPtrs)
                                       : crisoptix::Scene(ctx, meshes, sbt, emitters);

    // An invalid renderer size reaches the caller as crisoptix::Error with the library's message.
    {
        crisoptix::Camera c0({0, 0, 0}, {0, 0, -1}, {0, 1, 0}, 1.0f, 1.0f);
        bool threw = false;
        try {
            crisoptix::Renderer bad(0, H, ctx, c0, scene);
        } catch (const crisoptix::Error& err) {
            threw = err.Status() != CX_OK && std::string(err.what()).size() > 0;
            std::printf("expected error: %s\n", err.what());
        }
        if (!threw) throw std::runtime_error("a zero-width renderer was accepted");
    }

    const crisoptix::Vec3 camPos{camRec[0], camRec[1], camRec[2]}, camDir{camRec[3], camRec[4], camRec[5]},
        camUp{camRec[6], camRec[7], camRec[8]};
    crisoptix::Camera camA(camPos, camDir, camUp, 1.0f, camRec[9]), camB(camPos, camDir, camUp, 1.0f, camRec[9]);
    camA.SetAspectRatio(W, H);
    camB.SetAspectRatio(W, H);

    crisoptix::Renderer renA(W, H, ctx, camA, scene), renB(W, H, ctx, camB, scene);
    cx_launch_params p = renA.GetLaunchParams();
    for (const std::string& kv : overrides) set_field(p, kv);
    renA.SetLaunchParams(p);
    renB.SetLaunchParams(p);
    renA.Seed(seed);
    renB.Seed(seed);

    crisoptix::Pipeline pipe;
    uint32_t rg[CX_PASS_COUNT];
    for (int k = 0; k < CX_PASS_COUNT; k++) rg[k] = pipe.AddRaygenShader((cx_pass)k);
    pipe.CreatePipeline();

    const size_t px = (size_t)W * H * 3;
    float *outA = nullptr, *outB = nullptr;
    hip_check(hipMalloc(&outA, px * sizeof(float)), "hipMalloc");
    hip_check(hipMalloc(&outB, px * sizeof(float)), "hipMalloc");
    std::vector<float> hostA(px), hostB(px);
    std::ofstream out(argv[2], std::ios::binary);
    const float sceneSize = scene.GetSceneSize();
    const uint32_t nfr = (uint32_t)frames;
    out.write(reinterpret_cast<const char*>(&sceneSize), 4);
    out.write(reinterpret_cast<const char*>(&nfr), 4);
    const float kTurn = 2.0f * 3.14159265358979323846f / 64.0f;
    camA.SetChanged(true);
    camB.SetChanged(true);
    for (int f = 0; f < frames; f++) {
        if (orbit && f > 0) {
            camA.RotateAroundOrigin({0, 1, 0}, kTurn);
            camB.RotateAroundOrigin({0, 1, 0}, kTurn);
        }
        const cx_camera cd = camA.GetData();
        renA.LaunchFrame(outA);
        pipeline_frame(renB, pipe, rg, outB);
        hip_check(hipDeviceSynchronize(), "frame");
        hip_check(hipMemcpy(hostA.data(), outA, px * sizeof(float), hipMemcpyDeviceToHost), "copy");
        hip_check(hipMemcpy(hostB.data(), outB, px * sizeof(float), hipMemcpyDeviceToHost), "copy");
        const uint32_t fi = (uint32_t)renA.GetFrameIdx();
        out.write(reinterpret_cast<const char*>(&cd), sizeof(cd));
        out.write(reinterpret_cast<const char*>(&fi), 4);
        out.write(reinterpret_cast<const char*>(hostA.data()), (std::streamsize)(px * sizeof(float)));
        out.write(reinterpret_cast<const char*>(hostB.data()), (std::streamsize)(px * sizeof(float)));
        std::printf("frame %d: frameIdx %u, mean %.6f\n", f, fi, [&] {
            double s = 0;
            for (float v : hostA) s += v;
            return s / (double)px;
        }());
    }
    hip_check(hipFree(outA), "hipFree");
    hip_check(hipFree(outB), "hipFree");
    if (!out) throw std::runtime_error(std::string("cannot write ") + argv[2]);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "render_frames: %s\n", e.what());
        return 1;
    }
}
