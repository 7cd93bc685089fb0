// Drives the C++ drop-in classes (include/meshlib_amd/MRAmdDistanceField.h) the way MeshLib's
// callers do: meshToDistanceVolume's calcFromGridWithDistances with the voxel-centre basis,
// calcFromVector, calcSelfIntersections, the by-parts pipeline and findProjections.
// Input/output are raw little-endian files written/read by tests/test_gpu_adapter.py.
#include "meshlib_amd/MRAmdDistanceField.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

template <class T> static void rd(FILE* f, T* p, size_t n) { if (fread(p, sizeof(T), n, f) != n) { std::perror("read"); std::exit(2); } }
template <class T> static void wr(FILE* f, const T* p, size_t n) { fwrite(p, sizeof(T), n, f); }

int main(int argc, char** argv)
{
    if (argc != 3) { std::fprintf(stderr, "usage: adapter_main in.bin out.bin\n"); return 2; }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int32_t nv, nf, nq, dims[3];
    rd(in, &nv, 1); rd(in, &nf, 1); rd(in, &nq, 1); rd(in, dims, 3);
    MR::Mesh mesh;
    mesh.points.resize(nv);
    mesh.tris.resize(nf);
    rd(in, reinterpret_cast<float*>(mesh.points.data()), (size_t)nv * 3);
    rd(in, reinterpret_cast<int32_t*>(mesh.tris.data()), (size_t)nf * 3);
    std::vector<MR::Vector3f> pts(nq);
    rd(in, reinterpret_cast<float*>(pts.data()), (size_t)nq * 3);
    MR::AffineXf3f xf;
    rd(in, reinterpret_cast<float*>(&xf), 12);
    std::fclose(in);

    MR::Amd::FastWindingNumber fwn(mesh);
    const MR::Vector3i d{ dims[0], dims[1], dims[2] };
    std::vector<float> sdf, wn, sdfParts((size_t)dims[0] * dims[1] * dims[2], 0.f);
    if (auto r = fwn.calcFromGridWithDistances(sdf, d, xf, MR::DistanceToMeshOptions{}, {}); !r) {
        std::fprintf(stderr, "calcFromGridWithDistances: %s\n", r.error().c_str());
        return 1;
    }
    if (auto r = fwn.calcFromVector(wn, pts, 2.0f, {}, {}); !r) { std::fprintf(stderr, "%s\n", r.error().c_str()); return 1; }
    MR::FaceBitSet si;
    if (auto r = fwn.calcSelfIntersections(si, 2.0f, {}); !r) { std::fprintf(stderr, "%s\n", r.error().c_str()); return 1; }
    const size_t layer = (size_t)dims[0] * dims[1];
    auto r = fwn.calcFromGridWithDistancesByParts(
        [&](std::vector<float>&& data, const MR::Vector3i& pd, int z0) -> MR::Expected<void> {
            std::copy(data.begin(), data.end(), sdfParts.begin() + (size_t)z0 * layer);
            (void)pd;
            return {};
        },
        d, xf, MR::DistanceToMeshOptions{}, 1, {});
    if (!r) { std::fprintf(stderr, "byParts: %s\n", r.error().c_str()); return 1; }
    // cancellation surfaces as the canceled-operation error
    std::vector<float> tmp;
    auto c = fwn.calcFromGrid(tmp, d, xf, 2.0f, [](float p) { return p < 0.55f; });
    const int32_t canceled = (!c && c.error() == MR::stringOperationCanceled()) ? 1 : 0;

    MR::Amd::PointsToMeshProjector proj;
    proj.updateMeshData(&mesh);
    std::vector<MR::MeshProjectionResult> pr;
    proj.findProjections(pr, pts);

    FILE* out = std::fopen(argv[2], "wb");
    wr(out, sdf.data(), sdf.size());
    wr(out, sdfParts.data(), sdfParts.size());
    wr(out, wn.data(), wn.size());
    std::vector<uint8_t> sib(nf);
    for (int f = 0; f < nf; ++f) sib[f] = si.test(MR::FaceId(f)) ? 1 : 0;
    wr(out, sib.data(), sib.size());
    wr(out, reinterpret_cast<const uint8_t*>(pr.data()), pr.size() * 32);
    wr(out, &canceled, 1);
    std::fclose(out);
    std::printf("adapter ok\n");
    return 0;
}
