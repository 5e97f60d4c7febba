// ASCII VTU frame files in the layout the reference's solvers write
// (DEMSPH/kernel/DEMSolver.h:908-1098, WCSPHSolver.h:360-520): an
// UnstructuredGrid with TIME and STEP field data, one line per data array,
// every value preceded by a space, reals printed fixed with 10 decimals
// (std::fixed << std::setprecision(10), i.e. %.10f) from FP64 values.
// Host code only.
#pragma once

#include <cerrno>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include <sys/stat.h>

namespace demsph {
namespace vtu {

class Writer {
public:
    // dir/stem_FFFF.vtu (frame number zero-padded to at least 4 digits);
    // the directory is created when missing.
    Writer(const std::string& dir, const char* stem, long frame)
    {
        if (::mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST)
            throw std::runtime_error("cannot create directory " + dir);
        char name[64];
        std::snprintf(name, sizeof name, "/%s_%04ld.vtu", stem, frame);
        path_ = dir + name;
        f_ = std::fopen(path_.c_str(), "w");
        if (!f_) throw std::runtime_error("cannot open " + path_);
        std::setvbuf(f_, nullptr, _IOFBF, 1 << 20);
    }
    ~Writer()
    {
        if (f_) std::fclose(f_);
    }
    Writer(const Writer&) = delete;
    Writer& operator=(const Writer&) = delete;

    void begin(double time, long step, size_t points, size_t cells)
    {
        put("<?xml version=\"1.0\"?>\n"
            "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
            "  <UnstructuredGrid>\n"
            "    <FieldData>\n");
        std::fprintf(f_,
                     "      <DataArray type=\"Float32\" Name=\"TIME\"  NumberOfTuples=\"1\" format=\"ascii\"> %.10f "
                     "</DataArray>\n",
                     time);
        std::fprintf(f_,
                     "      <DataArray type=\"Int32\"   Name=\"STEP\"  NumberOfTuples=\"1\" format=\"ascii\"> %ld "
                     "</DataArray>\n",
                     step);
        put("    </FieldData>\n");
        std::fprintf(f_, "    <Piece NumberOfPoints=\"%zu\" NumberOfCells=\"%zu\">\n", points, cells);
    }

    // 3 values per point
    void points(const std::vector<double>& xyz)
    {
        put("      <Points>\n"
            "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n");
        for (double v : xyz) real(v);
        put("\n        </DataArray>\n"
            "      </Points>\n");
    }

    // one VTK_VERTEX cell per point
    void vertex_cells(size_t n)
    {
        put("      <Cells>\n"
            "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n");
        for (size_t i = 0; i < n; i++) std::fprintf(f_, " %zu", i);
        put("\n        </DataArray>\n"
            "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
        for (size_t i = 1; i <= n; i++) std::fprintf(f_, " %zu", i);
        put("\n        </DataArray>\n"
            "        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
        for (size_t i = 0; i < n; i++) put(" 1");
        put("\n        </DataArray>\n"
            "      </Cells>\n");
    }

    // one VTK_TRIANGLE cell per (t0, t1, t2)
    void triangle_cells(const std::vector<int>& t0, const std::vector<int>& t1, const std::vector<int>& t2)
    {
        const size_t n = t0.size();
        put("      <Cells>\n"
            "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n");
        for (size_t t = 0; t < n; t++) std::fprintf(f_, " %d %d %d", t0[t], t1[t], t2[t]);
        put("\n        </DataArray>\n"
            "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
        for (size_t t = 1; t <= n; t++) std::fprintf(f_, " %zu", 3 * t);
        put("\n        </DataArray>\n"
            "        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
        for (size_t t = 0; t < n; t++) put(" 5");
        put("\n        </DataArray>\n"
            "      </Cells>\n");
    }

    // "PointData" or "CellData", with the array named as the active scalars
    void open_data(const char* tag, const char* scalars) { std::fprintf(f_, "      <%s Scalars=\"%s\">\n", tag, scalars); }
    void close_data(const char* tag) { std::fprintf(f_, "      </%s>\n", tag); }

    void reals(const char* name, const std::vector<double>& v)
    {
        std::fprintf(f_, "        <DataArray type=\"Float32\" Name=\"%s\" format=\"ascii\">\n", name);
        for (double x : v) real(x);
        put("\n        </DataArray>\n");
    }
    // 3 values per tuple
    void vectors(const char* name, const std::vector<double>& v)
    {
        std::fprintf(f_, "        <DataArray type=\"Float32\" Name=\"%s\" NumberOfComponents=\"3\" format=\"ascii\">\n",
                     name);
        for (double x : v) real(x);
        put("\n        </DataArray>\n");
    }
    void ints(const char* name, const std::vector<int>& v)
    {
        std::fprintf(f_, "        <DataArray type=\"Int32\" Name=\"%s\" format=\"ascii\">\n", name);
        for (int x : v) std::fprintf(f_, " %d", x);
        put("\n        </DataArray>\n");
    }

    void end()
    {
        put("    </Piece>\n"
            "  </UnstructuredGrid>\n"
            "</VTKFile>\n");
        const bool bad = std::ferror(f_) != 0;
        const int rc = std::fclose(f_);
        f_ = nullptr;
        if (bad || rc != 0) throw std::runtime_error("write failed: " + path_);
    }

private:
    void put(const char* s) { std::fputs(s, f_); }
    void real(double v) { std::fprintf(f_, " %.10f", v); }
    std::FILE* f_ = nullptr;
    std::string path_;
};

// Background frame output: posted jobs run in order on one process-wide
// writer thread.  post() waits while two jobs are queued or running, which
// bounds the pinned snapshots held for them; drain() waits until every
// posted job has finished and returns the first error message since the
// last drain (empty if none).
void post(std::function<void()> job);
std::string drain();

}  // namespace vtu
}  // namespace demsph
