// MRMirror.h -- the few MeshLib types the adapter's signatures use, for builds without MeshLib.
//
// When MeshLib's own headers are available the adapter (MRAmdDistanceField.h) includes them
// instead and this file is not used.  Names, member layout and semantics follow:
//   Vector3 (source/MRMesh/MRVector3.h), Matrix3 / AffineXf3 (MRMatrix3.h, MRAffineXf.h),
//   Id (MRId.h: default -1, valid() >= 0), Expected / unexpected (MRExpected.h),
//   ProgressCallback (MRMeshFwd.h:748), DistanceToMeshOptions (MRDistanceToMeshOptions.h:10-33),
//   MeshProjectionResult (MRMeshProject.h:17-31), the interfaces IFastWindingNumber /
//   IFastWindingNumberByParts (MRFastWindingNumber.h:11-107) and IPointsToMeshProjector
//   (MRPointsToMeshProjector.h:12-30).
#pragma once

#include <array>
#include <cfloat>
#include <cstdint>
#include <functional>
#include <string>
#include <utility>
#include <vector>

namespace MR {

template <class T> struct Vector3 {
    T x{}, y{}, z{};
    Vector3() = default;
    Vector3(T x_, T y_, T z_) : x(x_), y(y_), z(z_) {}
};
using Vector3f = Vector3<float>;
using Vector3i = Vector3<int>;

struct Matrix3f {
    Vector3f x{ 1, 0, 0 }, y{ 0, 1, 0 }, z{ 0, 0, 1 };
};
struct AffineXf3f {
    Matrix3f A;
    Vector3f b;
    bool operator==(const AffineXf3f& o) const
    {
        const float* p = &A.x.x; const float* q = &o.A.x.x;
        for (int i = 0; i < 12; ++i) if (p[i] != q[i]) return false;
        return true;
    }
};
static_assert(sizeof(AffineXf3f) == 12 * sizeof(float), "AffineXf3f is A rows then b");

template <class Tag> struct Id {
    int id_ = -1;
    Id() = default;
    explicit Id(int i) : id_(i) {}
    bool valid() const { return id_ >= 0; }
    explicit operator bool() const { return id_ >= 0; }
    explicit operator int() const { return id_; }
};
struct FaceTag {}; struct EdgeTag {};
using FaceId = Id<FaceTag>;
using EdgeId = Id<EdgeTag>;

struct FaceBitSet {
    std::vector<bool> bits;
    void resize(size_t n) { bits.assign(n, false); }
    void set(FaceId f) { bits[(size_t)(int)f] = true; }
    bool test(FaceId f) const { return bits[(size_t)(int)f]; }
    size_t size() const { return bits.size(); }
};

using ProgressCallback = std::function<bool(float)>;

// tl::expected<void, std::string> in MeshLib
template <class T> class Expected;
template <> class Expected<void> {
public:
    Expected() = default;
    static Expected fail(std::string e) { Expected r; r.ok_ = false; r.err_ = std::move(e); return r; }
    bool has_value() const { return ok_; }
    explicit operator bool() const { return ok_; }
    const std::string& error() const { return err_; }
private:
    bool ok_ = true;
    std::string err_;
};
struct Unexpected { std::string e; operator Expected<void>() const { return Expected<void>::fail(e); } };
inline Unexpected unexpected(std::string e) { return { std::move(e) }; }
inline std::string stringOperationCanceled() { return "Operation was canceled"; }

struct DistanceToMeshOptions {
    float minDistSq{ 0 };
    float maxDistSq{ FLT_MAX };
    bool nullOutsideMinMax = true;
    float windingNumberThreshold = 0.5f;
    float windingNumberBeta = 2;
};

struct PointOnFace { FaceId face; Vector3f point; };
struct TriPointf { float a = 0, b = 0; };
struct MeshTriPoint { EdgeId e; TriPointf bary; };
struct MeshProjectionResult {
    PointOnFace proj;
    MeshTriPoint mtp;
    float distSq = 0;
    bool valid() const { return proj.face.valid(); }
};
static_assert(sizeof(MeshProjectionResult) == 32, "MeshProjectionResult wire record is 32 B");

// indexed triangle mesh as the adapter needs it: points and the triangulation in
// getTriangulation() order; edgeWithLeft is topology the mirror does not carry (-1)
struct Mesh {
    std::vector<Vector3f> points;
    std::vector<std::array<int, 3>> tris;
    EdgeId edgeWithLeft(FaceId) const { return EdgeId(); }
};

class IFastWindingNumber {
public:
    virtual ~IFastWindingNumber() = default;
    virtual Expected<void> calcFromVector(std::vector<float>& res, const std::vector<Vector3f>& points, float beta,
                                          FaceId skipFace = {}, const ProgressCallback& cb = {}) = 0;
    virtual Expected<void> calcSelfIntersections(FaceBitSet& res, float beta, const ProgressCallback& cb = {}) = 0;
    virtual Expected<void> calcFromGrid(std::vector<float>& res, const Vector3i& dims, const AffineXf3f& gridToMeshXf,
                                        float beta, const ProgressCallback& cb = {}) = 0;
    virtual Expected<void> calcFromGridWithDistances(std::vector<float>& res, const Vector3i& dims,
                                                     const AffineXf3f& gridToMeshXf,
                                                     const DistanceToMeshOptions& options,
                                                     const ProgressCallback& cb) = 0;
};

class IFastWindingNumberByParts {
public:
    virtual ~IFastWindingNumberByParts() = default;
    using GridByPartsFunc = std::function<Expected<void>(std::vector<float>&& data, const Vector3i& dims, int zOffset)>;
    virtual Expected<void> calcFromGridByParts(GridByPartsFunc resFunc, const Vector3i& dims,
                                               const AffineXf3f& gridToMeshXf, float beta, int layerOverlap,
                                               const ProgressCallback& cb) = 0;
    virtual Expected<void> calcFromGridWithDistancesByParts(GridByPartsFunc resFunc, const Vector3i& dims,
                                                            const AffineXf3f& gridToMeshXf,
                                                            const DistanceToMeshOptions& options, int layerOverlap,
                                                            const ProgressCallback& cb) = 0;
};

class IPointsToMeshProjector {
public:
    virtual ~IPointsToMeshProjector() = default;
    virtual void updateMeshData(const Mesh* mesh) = 0;
    virtual void findProjections(std::vector<MeshProjectionResult>& result, const std::vector<Vector3f>& points,
                                 const AffineXf3f* worldXf = nullptr, const AffineXf3f* worldRefXf = nullptr,
                                 float upDistLimitSq = FLT_MAX, float loDistLimitSq = 0.0f) = 0;
    virtual size_t projectionsHeapBytes(size_t numProjections) const = 0;
};

} // namespace MR
