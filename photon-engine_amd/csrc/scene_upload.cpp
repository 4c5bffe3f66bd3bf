// scene_upload.cpp — flattens a pe_scene_desc (Scene::shapes() order) into the device layout:
// vertex arrays, per-triangle shading records in scene order, the BVH and its leaf-ordered TriAccel array.
//
// Compiled without floating-point contraction: the per-triangle threshold 1e-14 |e1|^2 |e2|^2 and the edges
// must be the same floats the reference computes (geometry/bvh.cpp:45-50, triangle.cpp:12-22).
#include "pe_scene_host.h"
#include "bvh_build.h"
#include "pe_traverse.h"

#include <chrono>
#include <cstring>
#include <cmath>
#include <string>

namespace pe {

namespace {

bool finite_vec(const float* p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

}  // namespace

int upload_scene(pe_device* dev, const pe_scene_desc& desc, pe_scene& sc) {
    if (desc.num_meshes < 0 || (desc.num_meshes > 0 && !desc.meshes)) {
        set_error("scene: bad mesh list");
        return PE_ERR_INVALID;
    }
    if (desc.num_materials < 0 || (desc.num_materials > 0 && !desc.materials)) {
        set_error("scene: bad material list");
        return PE_ERR_INVALID;
    }
    if (desc.num_lights < 0 || (desc.num_lights > 0 && !desc.lights)) {
        set_error("scene: bad light list");
        return PE_ERR_INVALID;
    }
    uint64_t nv = 0, nt = 0;
    for (int s = 0; s < desc.num_meshes; ++s) {
        const pe_mesh_desc& m = desc.meshes[s];
        if (m.num_triangles > 0 && (!m.positions || !m.indices)) {
            set_error("scene: mesh " + std::to_string(s) + " has no positions or indices");
            return PE_ERR_INVALID;
        }
        if (m.material < 0 || m.material >= desc.num_materials) {
            set_error("scene: mesh " + std::to_string(s) + " material index out of range");
            return PE_ERR_INVALID;
        }
        if (m.emitter_base >= desc.num_lights || (m.emitter_base >= 0 && !m.standalone &&
                                                  static_cast<uint64_t>(m.emitter_base) + m.num_triangles >
                                                      static_cast<uint64_t>(desc.num_lights))) {
            set_error("scene: mesh " + std::to_string(s) + " emitter range outside the light list");
            return PE_ERR_INVALID;
        }
        for (uint64_t k = 0; k < 3ull * m.num_triangles; ++k)
            if (m.indices[k] >= m.num_vertices) {
                set_error("scene: mesh " + std::to_string(s) + " index " + std::to_string(m.indices[k]) +
                          " >= num_vertices " + std::to_string(m.num_vertices));
                return PE_ERR_INVALID;
            }
        if (!finite_vec(m.positions, 3ull * m.num_vertices)) {
            set_error("scene: mesh " + std::to_string(s) + " has non-finite positions");
            return PE_ERR_INVALID;
        }
        nv += m.num_vertices;
        nt += m.num_triangles;
        sc.max_material_ref = std::max(sc.max_material_ref, m.material);
        if (m.emitter_base >= 0) {
            const uint64_t last = m.standalone || m.num_triangles == 0 ? static_cast<uint64_t>(m.emitter_base)
                                                                        : m.emitter_base + m.num_triangles - 1;
            sc.max_emitter_ref = std::max(sc.max_emitter_ref, static_cast<int>(last));
        }
    }
    if (nt > kMaxTraceTriangles) {
        set_error("scene: more than 2^27 triangles (the traversal's leaf items address 2^27)");
        return PE_ERR_UNSUPPORTED;
    }

    std::vector<float4> pos(nv), nrm(nv);
    std::vector<float2> uv(nv);
    std::vector<TriShade> shade(nt);
    std::vector<float> lo(3 * nt), hi(3 * nt);
    uint64_t vb = 0, tb = 0;
    for (int s = 0; s < desc.num_meshes; ++s) {
        const pe_mesh_desc& m = desc.meshes[s];
        for (uint32_t v = 0; v < m.num_vertices; ++v) {
            pos[vb + v] = make_float4(m.positions[3 * v], m.positions[3 * v + 1], m.positions[3 * v + 2], 0.0f);
            nrm[vb + v] = m.normals ? make_float4(m.normals[3 * v], m.normals[3 * v + 1], m.normals[3 * v + 2], 0.0f)
                                    : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            uv[vb + v] = m.uvs ? make_float2(m.uvs[2 * v], m.uvs[2 * v + 1]) : make_float2(0.0f, 0.0f);
        }
        // A TriangleMesh given without normals gets area-weighted vertex normals (TriangleMesh ctor and
        // calculateNormals, mesh.cpp:22-54): face normals (v1 - v0) x (v2 - v0) summed per vertex in triangle
        // order, then normalised. Standalone triangles keep what they were given.
        const bool computed = !m.normals && !m.standalone && m.num_vertices > 0 && m.num_triangles > 0;
        if (computed) {
            std::vector<V3> acc(m.num_vertices, mk(0.0f, 0.0f, 0.0f));
            auto P = [&](uint32_t i) { return mk(m.positions[3 * i], m.positions[3 * i + 1], m.positions[3 * i + 2]); };
            for (uint32_t t = 0; t < m.num_triangles; ++t) {
                const uint32_t i0 = m.indices[3 * t], i1 = m.indices[3 * t + 1], i2 = m.indices[3 * t + 2];
                const V3 n = cross(P(i1) - P(i0), P(i2) - P(i0));
                acc[i0] = acc[i0] + n;
                acc[i1] = acc[i1] + n;
                acc[i2] = acc[i2] + n;
            }
            for (uint32_t v = 0; v < m.num_vertices; ++v) {
                V3 n = acc[v];
                if (dot(n, n) > 0.0f) {
                    const float len = std::sqrt(dot(n, n));
                    if (len <= 1e-20f) {
                        n = mk(0.0f, 0.0f, 0.0f);
                    } else {
                        const float inv = 1.0f / len;
                        n = mk(n.x * inv, n.y * inv, n.z * inv);
                    }
                }
                nrm[vb + v] = make_float4(n.x, n.y, n.z, 0.0f);
            }
        }
        for (uint32_t t = 0; t < m.num_triangles; ++t) {
            TriShade& r = shade[tb + t];
            r.v0 = static_cast<uint32_t>(vb + m.indices[3 * t]);
            r.v1 = static_cast<uint32_t>(vb + m.indices[3 * t + 1]);
            r.v2 = static_cast<uint32_t>(vb + m.indices[3 * t + 2]);
            r.material = m.material;
            r.shape = s;
            r.prim = m.standalone ? -1 : static_cast<int32_t>(t);
            r.emitter = m.emitter_base < 0 ? -1 : (m.standalone ? m.emitter_base : m.emitter_base + static_cast<int32_t>(t));
            r.flags = ((m.normals || computed) ? kTriHasNormals : 0u) | (m.uvs ? kTriHasUv : 0u) |
                      ((m.flags & PE_SHAPE_CAMERA_INVISIBLE) ? kTriCameraInvisible : 0u);
            const float4 p[3] = {pos[r.v0], pos[r.v1], pos[r.v2]};
            for (int k = 0; k < 3; ++k) {
                const float a = k == 0 ? p[0].x : (k == 1 ? p[0].y : p[0].z);
                const float b = k == 0 ? p[1].x : (k == 1 ? p[1].y : p[1].z);
                const float c = k == 0 ? p[2].x : (k == 1 ? p[2].y : p[2].z);
                lo[3 * (tb + t) + k] = std::fmin(a, std::fmin(b, c));
                hi[3 * (tb + t) + k] = std::fmax(a, std::fmax(b, c));
            }
        }
        vb += m.num_vertices;
        tb += m.num_triangles;
    }

    BvhBuildOptions opt;
    const BvhBuildResult bvh = build_bvh(lo.data(), hi.data(), nt, opt);
    int depth4 = 0;
    std::vector<BvhNode4> nodes4 = collapse_bvh4(bvh.nodes, depth4);
    if (nt == 0) {
        // A scene of no triangles (an environment alone): one root whose four slots are empty boxes, so every
        // ray misses; the shading buffers get one zero record each so that no device pointer is null.
        BvhNode4 root{};
        root.lox = root.loy = root.loz = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        root.hix = root.hiy = root.hiz = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
        root.item = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
        nodes4.assign(1, root);
        depth4 = 1;
        shade.resize(1);
        shade[0].material = 0;
        shade[0].emitter = -1;
        if (pos.empty()) {
            pos.resize(1);
            nrm.resize(1);
            uv.resize(1);
        }
    }

    // One zero record of tail padding: traversal reads every triangle as a 64-byte item.
    std::vector<TriAccel> acc(nt + 1);
    acc[nt] = TriAccel{};
    for (uint64_t i = 0; i < nt; ++i) {
        const uint32_t t = bvh.order[i];
        const TriShade& r = shade[t];
        const float4 p0 = pos[r.v0], p1 = pos[r.v1], p2 = pos[r.v2];
        const float e1x = p1.x - p0.x, e1y = p1.y - p0.y, e1z = p1.z - p0.z;
        const float e2x = p2.x - p0.x, e2y = p2.y - p0.y, e2z = p2.z - p0.z;
        const float l1 = e1x * e1x + e1y * e1y + e1z * e1z;
        const float l2 = e2x * e2x + e2y * e2y + e2z * e2z;
        const float eps = 1e-14f * l1 * l2;
        int32_t id = static_cast<int32_t>(t);
        float idf;
        std::memcpy(&idf, &id, 4);
        acc[i].v0 = make_float4(p0.x, p0.y, p0.z, eps);
        acc[i].e1 = make_float4(e1x, e1y, e1z, idf);
        acc[i].e2 = make_float4(e2x, e2y, e2z, 0.0f);
    }

    hipStream_t s = dev->stream;
    if (!sc.nodes.upload(nodes4.data(), nodes4.size() * sizeof(BvhNode4), s) ||
        !sc.tris.upload(acc.data(), acc.size() * sizeof(TriAccel), s) ||
        !sc.tri_shade.upload(shade.data(), shade.size() * sizeof(TriShade), s) ||
        !sc.positions.upload(pos.data(), pos.size() * sizeof(float4), s) ||
        !sc.normals.upload(nrm.data(), nrm.size() * sizeof(float4), s) ||
        !sc.uvs.upload(uv.data(), uv.size() * sizeof(float2), s))
        return PE_ERR_OOM;
    // The host vectors die at return: wait for the copies that read them.
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        set_error(std::string("scene upload: ") + hipGetErrorString(e));
        return PE_ERR_HIP;
    }
    sc.num_nodes = nodes4.size();
    const int rc = prepare_shading(dev, desc, sc);
    if (rc != PE_OK) return rc;
    sc.num_tris = nt;
    sc.num_verts = nv;
    sc.num_shapes = desc.num_meshes;
    sc.bvh_depth = depth4;
    return PE_OK;
}

}  // namespace pe
