// MRAmdDistanceField.h -- C++ drop-ins for MeshLib's GPU distance-field classes on MI355X.
//
//   MR::Amd::FastWindingNumber     replaces MR::Cuda::FastWindingNumber
//                                  (source/MRCuda/MRCudaFastWindingNumber.h:17-38)
//   MR::Amd::PointsToMeshProjector replaces MR::Cuda::PointsToMeshProjector
//                                  (source/MRCuda/MRCudaPointsToMeshProjector.h:9-31)
//
// Header-only over the C-ABI of meshlib_amd.h.  With MeshLib's headers on the include path the
// classes implement MeshLib's own interfaces (IFastWindingNumber, IFastWindingNumberByParts,
// IPointsToMeshProjector) and take MR::Mesh; without them they implement the signature-identical
// mirror in MRMirror.h.  Semantics follow the reference GPU classes: the device data is built on
// first use (prepareData_), results are resized by the callee, errors come back as
// Expected<void> ("AMD GPU error: ..." or the canceled-operation string), and the projector's
// void interface only logs errors to stderr.
#pragma once

#include "../meshlib_amd.h"

#if defined(__has_include)
#if __has_include(<MRMesh/MRFastWindingNumber.h>) && !defined(MESHLIB_AMD_USE_MIRROR)
#define MESHLIB_AMD_WITH_MESHLIB 1
#endif
#endif

#ifdef MESHLIB_AMD_WITH_MESHLIB
#include <MRMesh/MRDistanceToMeshOptions.h>
#include <MRMesh/MRFastWindingNumber.h>
#include <MRMesh/MRMesh.h>
#include <MRMesh/MRMeshProject.h>
#include <MRMesh/MRPointsToMeshProjector.h>
#else
#include "MRMirror.h"
#endif

#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

namespace MR {
namespace Amd {

namespace detail {

inline std::string lastError()
{
    char buf[1024];
    mi_last_error(buf, sizeof(buf));
    return buf;
}

inline Expected<void> toExpected(int rc)
{
    if (rc == MI_OK) return {};
    if (rc == MI_E_CANCELED) return unexpected(stringOperationCanceled());
    return unexpected(lastError());
}

inline const float* xfPtr(const AffineXf3f& xf) { return reinterpret_cast<const float*>(&xf); }

#ifdef MESHLIB_AMD_WITH_MESHLIB
inline const float* verts(const Mesh& m) { return reinterpret_cast<const float*>(m.points.vec_.data()); }
inline int32_t numVerts(const Mesh& m) { return (int32_t)m.points.size(); }
inline EdgeId edgeWithLeft(const Mesh& m, FaceId f) { return m.topology.edgeWithLeft(f); }
struct Triangles {
    Triangulation t;
    explicit Triangles(const Mesh& m) : t(m.topology.getTriangulation()) {}
    const int32_t* data() const { return reinterpret_cast<const int32_t*>(t.vec_.data()); }
    int32_t size() const { return (int32_t)t.size(); }
};
#else
inline const float* verts(const Mesh& m) { return reinterpret_cast<const float*>(m.points.data()); }
inline int32_t numVerts(const Mesh& m) { return (int32_t)m.points.size(); }
inline EdgeId edgeWithLeft(const Mesh& m, FaceId f) { return m.edgeWithLeft(f); }
struct Triangles {
    const Mesh& m;
    explicit Triangles(const Mesh& mm) : m(mm) {}
    const int32_t* data() const { return reinterpret_cast<const int32_t*>(m.tris.data()); }
    int32_t size() const { return (int32_t)m.tris.size(); }
};
#endif

// device-resident mesh data (tree, dipoles, triangles), shared by both classes
struct MeshData {
    mi_mesh* h = nullptr;
    ~MeshData() { if (h) mi_mesh_destroy(h); }
    static std::shared_ptr<MeshData> make(const Mesh& mesh, std::string* err)
    {
        auto d = std::make_shared<MeshData>();
        Triangles t(mesh);
        const int rc = mi_mesh_create(verts(mesh), numVerts(mesh), t.data(), t.size(), &d->h);
        if (rc != MI_OK) {
            if (err) *err = lastError();
            return nullptr;
        }
        return d;
    }
};

struct CbThunk {
    const ProgressCallback* cb;
    static int call(float v, void* u)
    {
        const auto* t = static_cast<const CbThunk*>(u);
        return (!t->cb || !*t->cb || (*t->cb)(v)) ? 1 : 0;
    }
};

struct PartThunk {
    IFastWindingNumberByParts::GridByPartsFunc* fn;
    std::string err;
    static int call(const float* data, const int32_t dims[3], int32_t zOffset, void* u)
    {
        auto* t = static_cast<PartThunk*>(u);
        std::vector<float> v(data, data + (size_t)dims[0] * dims[1] * dims[2]);
        auto r = (*t->fn)(std::move(v), Vector3i{ dims[0], dims[1], dims[2] }, zOffset);
        if (!r) { t->err = r.error(); return MI_E_CANCELED; }
        return MI_OK;
    }
};

inline mi_dist_options toOpts(const DistanceToMeshOptions& o)
{
    return { o.minDistSq, o.maxDistSq, o.nullOutsideMinMax ? 1 : 0, o.windingNumberThreshold,
             o.windingNumberBeta };
}

} // namespace detail

/// GPU winding numbers and signed distances on MI355X; drop-in for MR::Cuda::FastWindingNumber
class FastWindingNumber : public IFastWindingNumber, public IFastWindingNumberByParts {
    const Mesh& mesh_;
    std::shared_ptr<detail::MeshData> data_;
    std::mutex mu_;

    Expected<void> prepareData_(const ProgressCallback& cb)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (data_) return {};
        if (cb && !cb(0.0f)) return unexpected(stringOperationCanceled());
        std::string err;
        data_ = detail::MeshData::make(mesh_, &err);
        if (!data_) return unexpected(err);
        return {};
    }

public:
    explicit FastWindingNumber(const Mesh& mesh) : mesh_(mesh) {}

    Expected<void> calcFromVector(std::vector<float>& res, const std::vector<Vector3f>& points, float beta,
                                  FaceId skipFace = {}, const ProgressCallback& cb = {}) override
    {
        if (auto p = prepareData_(cb); !p) return p;
        res.resize(points.size());
        detail::CbThunk t{ &cb };
        return detail::toExpected(mi_fwn_calc_from_vector(data_->h, reinterpret_cast<const float*>(points.data()),
                                                          (int64_t)points.size(), beta,
                                                          skipFace.valid() ? (int)skipFace : -1, res.data(),
                                                          &detail::CbThunk::call, &t));
    }

    Expected<void> calcSelfIntersections(FaceBitSet& res, float beta, const ProgressCallback& cb = {}) override
    {
        if (auto p = prepareData_(cb); !p) return p;
        detail::Triangles tr(mesh_);
        std::vector<uint8_t> flags((size_t)tr.size());
        detail::CbThunk t{ &cb };
        const int rc = mi_fwn_calc_self_intersections(data_->h, beta, flags.data(), &detail::CbThunk::call, &t);
        if (rc != MI_OK) return detail::toExpected(rc);
        res.resize(flags.size());
        for (size_t f = 0; f < flags.size(); ++f)
            if (flags[f]) res.set(FaceId((int)f));
        return {};
    }

    Expected<void> calcFromGrid(std::vector<float>& res, const Vector3i& dims, const AffineXf3f& gridToMeshXf,
                                float beta, const ProgressCallback& cb = {}) override
    {
        if (auto p = prepareData_(cb); !p) return p;
        res.resize((size_t)dims.x * dims.y * dims.z);
        const int32_t d[3] = { dims.x, dims.y, dims.z };
        detail::CbThunk t{ &cb };
        return detail::toExpected(mi_fwn_calc_from_grid(data_->h, d, detail::xfPtr(gridToMeshXf), beta, res.data(),
                                                        &detail::CbThunk::call, &t));
    }

    Expected<void> calcFromGridWithDistances(std::vector<float>& res, const Vector3i& dims,
                                             const AffineXf3f& gridToMeshXf, const DistanceToMeshOptions& options,
                                             const ProgressCallback& cb) override
    {
        if (auto p = prepareData_(cb); !p) return p;
        res.resize((size_t)dims.x * dims.y * dims.z);
        const int32_t d[3] = { dims.x, dims.y, dims.z };
        const mi_dist_options o = detail::toOpts(options);
        detail::CbThunk t{ &cb };
        return detail::toExpected(mi_fwn_calc_from_grid_with_distances(data_->h, d, detail::xfPtr(gridToMeshXf), &o,
                                                                       res.data(), &detail::CbThunk::call, &t));
    }

    Expected<void> calcFromGridByParts(GridByPartsFunc resFunc, const Vector3i& dims, const AffineXf3f& gridToMeshXf,
                                       float beta, int layerOverlap, const ProgressCallback& cb) override
    {
        if (auto p = prepareData_(cb); !p) return p;
        const int32_t d[3] = { dims.x, dims.y, dims.z };
        detail::PartThunk pt{ &resFunc, {} };
        detail::CbThunk t{ &cb };
        const int rc = mi_fwn_calc_from_grid_by_parts(data_->h, &detail::PartThunk::call, &pt, d,
                                                      detail::xfPtr(gridToMeshXf), beta, layerOverlap,
                                                      &detail::CbThunk::call, &t);
        if (rc != MI_OK && !pt.err.empty()) return unexpected(pt.err);
        return detail::toExpected(rc);
    }

    Expected<void> calcFromGridWithDistancesByParts(GridByPartsFunc resFunc, const Vector3i& dims,
                                                    const AffineXf3f& gridToMeshXf,
                                                    const DistanceToMeshOptions& options, int layerOverlap,
                                                    const ProgressCallback& cb) override
    {
        if (auto p = prepareData_(cb); !p) return p;
        const int32_t d[3] = { dims.x, dims.y, dims.z };
        const mi_dist_options o = detail::toOpts(options);
        detail::PartThunk pt{ &resFunc, {} };
        detail::CbThunk t{ &cb };
        const int rc = mi_fwn_calc_from_grid_with_distances_by_parts(data_->h, &detail::PartThunk::call, &pt, d,
                                                                     detail::xfPtr(gridToMeshXf), &o, layerOverlap,
                                                                     &detail::CbThunk::call, &t);
        if (rc != MI_OK && !pt.err.empty()) return unexpected(pt.err);
        return detail::toExpected(rc);
    }
};

/// GPU closest points on MI355X; drop-in for MR::Cuda::PointsToMeshProjector
class PointsToMeshProjector : public IPointsToMeshProjector {
    std::shared_ptr<detail::MeshData> data_;
    const Mesh* mesh_ = nullptr;

public:
    void updateMeshData(const Mesh* mesh) override
    {
        mesh_ = mesh;
        data_.reset();
        if (!mesh) return;
        std::string err;
        data_ = detail::MeshData::make(*mesh, &err);
        if (!data_) std::fprintf(stderr, "PointsToMeshProjector: %s\n", err.c_str());
    }

    void findProjections(std::vector<MeshProjectionResult>& result, const std::vector<Vector3f>& points,
                         const AffineXf3f* objXf = nullptr, const AffineXf3f* refObjXf = nullptr,
                         float upDistLimitSq = FLT_MAX, float loDistLimitSq = 0.0f) override
    {
        if (!mesh_ || !data_) return;
        result.resize(points.size());
        static_assert(sizeof(MeshProjectionResult) == sizeof(mi_projection), "32 B projection record");
        const int rc = mi_proj_find_projections(
            data_->h, reinterpret_cast<const float*>(points.data()), (int64_t)points.size(),
            objXf ? detail::xfPtr(*objXf) : nullptr, refObjXf ? detail::xfPtr(*refObjXf) : nullptr, upDistLimitSq,
            loDistLimitSq, reinterpret_cast<mi_projection*>(result.data()));
        if (rc != MI_OK) {
            std::fprintf(stderr, "PointsToMeshProjector: %s\n", detail::lastError().c_str());
            return;
        }
        // the kernels write e = -1; edgeWithLeft is topology (MRCudaPointsToMeshProjector.cpp:118-124)
        for (auto& r : result)
            if (r.proj.face.valid()) r.mtp.e = detail::edgeWithLeft(*mesh_, r.proj.face);
    }

    size_t projectionsHeapBytes(size_t numProjections) const override
    {
        return mi_proj_heap_bytes(numProjections);
    }
};

} // namespace Amd
} // namespace MR
