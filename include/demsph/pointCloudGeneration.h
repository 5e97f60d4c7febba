// Scene builders the reference's applications call: uniform lattices, hex
// sheets with optional periodic snapping, and triangulated cylinders.
//
// Reference: DEMSPH/pointCloudGeneration.h (generateUniformPointCloud :7-56,
// HexPBC2D / makeHexPBC2D / generateHexPointCloud :60-195, TriangleMesh /
// makeCylinderMesh :197-327).  Same names and arguments; the outputs are the
// same point sets in the same order (pinned in tests against the
// reference's own frames: SPH_0000.vtu's 18,081 lattice points and
// triangleWall_0000.vtu's 1,334 vertices / 2,664 triangles).
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

#include "demsph_types.h"

// Inclusive lattice from minBound to maxBound at spacing dx, x fastest, then
// y, then z; floor(length/dx + 1e-12) + 1 points per axis.
inline std::vector<double3> generateUniformPointCloud(const double3 minBound, const double3 maxBound,
                                                      const double spacing)
{
    std::vector<double3> pts;
    if (spacing <= 0.0) return pts;
    const double eps = 1e-12;
    const double lo[3] = {std::min(minBound.x, maxBound.x), std::min(minBound.y, maxBound.y),
                          std::min(minBound.z, maxBound.z)};
    const double hi[3] = {std::max(minBound.x, maxBound.x), std::max(minBound.y, maxBound.y),
                          std::max(minBound.z, maxBound.z)};
    std::vector<double> axis[3];
    for (int a = 0; a < 3; a++) {
        const int count = std::max(0, (int)std::floor((hi[a] - lo[a]) / spacing + eps) + 1);
        for (int k = 0; k < count; k++) {
            const double v = lo[a] + k * spacing;
            if (v <= hi[a] + eps) axis[a].push_back(v);
        }
    }
    pts.reserve(axis[0].size() * axis[1].size() * axis[2].size());
    for (double z : axis[2])
        for (double y : axis[1])
            for (double x : axis[0]) pts.push_back(make_double3(x, y, z));
    return pts;
}

// Periodic cell of a hex sheet in the x-y plane: rows dy = a sqrt(3)/2
// apart, odd rows shifted by a/2.  A periodic axis is snapped to a whole
// number of steps; wrapping y by Ly also shifts x by shearX.
struct HexPBC2D {
    double3 min{0.0, 0.0, 0.0};
    double3 max{0.0, 0.0, 0.0};
    double Lx{0.0};
    double Ly{0.0};
    double shearX{0.0};
    int nx{0};
    int ny{0};
    bool periodicX{false};
    bool periodicY{false};
};

inline double ceilDivLen(const double len, const double step)
{
    if (step <= 0.0) return 0.0;
    return std::ceil(len / step - 1e-12);
}

inline HexPBC2D makeHexPBC2D(const double3 minBound, const double3 maxBound, const double spacing,
                             const bool periodicX, const bool periodicY)
{
    HexPBC2D c;
    c.periodicX = periodicX;
    c.periodicY = periodicY;
    if (spacing <= 0.0) return c;
    const double ax = spacing, ay = spacing * std::sqrt(3.0) * 0.5;
    const double lx = std::max(0.0, maxBound.x - minBound.x), ly = std::max(0.0, maxBound.y - minBound.y);
    c.nx = periodicX ? std::max(1, (int)ceilDivLen(lx, ax)) : std::max(1, (int)std::floor(lx / ax + 1e-12));
    c.ny = periodicY ? std::max(1, (int)ceilDivLen(ly, ay)) : std::max(1, (int)std::floor(ly / ay + 1e-12));
    c.min = minBound;
    c.Lx = c.nx * ax;
    c.Ly = c.ny * ay;
    c.max = make_double3(periodicX ? minBound.x + c.Lx : maxBound.x, periodicY ? minBound.y + c.Ly : maxBound.y,
                         maxBound.z);
    c.shearX = (periodicY && (c.ny & 1)) ? 0.5 * ax : 0.0;
    return c;
}

// Hex-packed points at z = minBound.z, rows in y, points along x within a
// row.  Non-periodic axes stop before maxBound.
inline void generateHexPointCloud(const double3 minBound, const double3 maxBound, const double spacing,
                                  const bool periodicX, const bool periodicY, std::vector<double3>& points,
                                  HexPBC2D& pbc)
{
    points.clear();
    pbc = makeHexPBC2D(minBound, maxBound, spacing, periodicX, periodicY);
    if (pbc.nx <= 0 || pbc.ny <= 0 || spacing <= 0.0) return;
    const double ax = spacing, ay = spacing * std::sqrt(3.0) * 0.5;
    points.reserve((size_t)pbc.nx * (size_t)pbc.ny);
    for (int j = 0; j < pbc.ny; j++) {
        const double y = pbc.min.y + j * ay;
        if (!periodicY && y >= maxBound.y) break;
        const double shift = (j & 1) ? 0.5 * ax : 0.0;
        for (int i = 0; i < pbc.nx; i++) {
            const double x = pbc.min.x + i * ax + shift;
            if (!periodicX && x >= maxBound.x) break;
            points.push_back(make_double3(x, y, pbc.min.z));
        }
    }
}

struct TriangleMesh {
    std::vector<double3> vertices;
    std::vector<int> tri0, tri1, tri2;
    void reserveTriangles(size_t n)
    {
        tri0.reserve(n);
        tri1.reserve(n);
        tri2.reserve(n);
    }
    void addTriangle(int i0, int i1, int i2)
    {
        tri0.push_back(i0);
        tri1.push_back(i1);
        tri2.push_back(i2);
    }
    size_t numTriangles() const { return tri0.size(); }
};

namespace demsph {
namespace mesh_detail {
inline double3 cross3(double3 a, double3 b)
{
    return make_double3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double3 unit3(double3 a)
{
    const double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return l > 0.0 ? make_double3(a.x / l, a.y / l, a.z / l) : a;
}
}  // namespace mesh_detail
}  // namespace demsph

// Side of a cylinder as nHeight+1 rings of nAround vertices, two triangles
// per quad (i00,i10,i11) and (i00,i11,i01), plus optional cap fans around a
// centre vertex at each end.
inline TriangleMesh makeCylinderMesh(const double3& center, const double3& axis, double radius, double height,
                                     int nAround, int nHeight, bool caps = true)
{
    using namespace demsph::mesh_detail;
    TriangleMesh m;
    if (nAround < 3 || nHeight < 1) return m;
    const double3 w = unit3(axis);
    const double3 tmp = std::fabs(w.z) < 0.9 ? make_double3(0, 0, 1) : make_double3(0, 1, 0);
    const double3 u = unit3(cross3(tmp, w));
    const double3 v = unit3(cross3(w, u));
    for (int iz = 0; iz <= nHeight; iz++) {
        const double z = ((double)iz / nHeight - 0.5) * height;
        const double3 c = make_double3(center.x + z * w.x, center.y + z * w.y, center.z + z * w.z);
        for (int ia = 0; ia < nAround; ia++) {
            const double ang = 2.0 * pi() * ((double)ia / nAround);
            const double ca = std::cos(ang), sa = std::sin(ang);
            m.vertices.push_back(make_double3(c.x + radius * (ca * u.x + sa * v.x), c.y + radius * (ca * u.y + sa * v.y),
                                              c.z + radius * (ca * u.z + sa * v.z)));
        }
    }
    auto vid = [&](int iz, int ia) { return iz * nAround + (ia % nAround); };
    m.reserveTriangles((size_t)2 * nAround * nHeight + (caps ? 2 * (size_t)nAround : 0));
    for (int iz = 0; iz < nHeight; iz++)
        for (int ia = 0; ia < nAround; ia++) {
            const int i00 = vid(iz, ia), i01 = vid(iz, ia + 1), i10 = vid(iz + 1, ia), i11 = vid(iz + 1, ia + 1);
            m.addTriangle(i00, i10, i11);
            m.addTriangle(i00, i11, i01);
        }
    if (caps) {
        const int bc = (int)m.vertices.size();
        m.vertices.push_back(make_double3(center.x - 0.5 * height * w.x, center.y - 0.5 * height * w.y,
                                          center.z - 0.5 * height * w.z));
        const int tc = (int)m.vertices.size();
        m.vertices.push_back(make_double3(center.x + 0.5 * height * w.x, center.y + 0.5 * height * w.y,
                                          center.z + 0.5 * height * w.z));
        for (int ia = 0; ia < nAround; ia++) m.addTriangle(bc, vid(0, ia + 1), vid(0, ia));
        for (int ia = 0; ia < nAround; ia++) m.addTriangle(tc, vid(nHeight, ia), vid(nHeight, ia + 1));
    }
    return m;
}
