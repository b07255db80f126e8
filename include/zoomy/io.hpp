// Snapshot output of a run: the reference's IOManager (IOManager.hpp:19-120, 346-376) for the
// VTK path.  Snapshots every t_end / io.snapshots from time 0 (ShouldWrite / AdvanceSnapshot
// with the same tolerance), files "<filename>-<idx:03d>.vtu" in io.directory plus the
// ParaView series file "<filename>.vtu.series" with the same layout.  The .vtu is VTK XML
// UnstructuredGrid with appended raw binary data: the mesh vertices, the cells in the mesh
// file's order, and two cell arrays, "State" (n_q components) and "Aux" (n_aux components),
// the two fields of the reference's output DM (VirtualSolver.hpp:482-497, PackState 359-384).
// PETSc's own array naming inside the file is not reproduced.
#pragma once
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "zdm_mesh.h"
#include "zoomy/settings.hpp"

namespace zoomy {

class IOManager {
public:
    explicit IOManager(const Settings& s) : settings_(s) {
        dt_snap_ = s.io.snapshots > 0 ? s.solver.t_end / s.io.snapshots : 0.0;
    }

    bool enabled() const { return settings_.io.snapshots > 0; }

    // IOManager::PrepareDirectory: create io.directory, or empty it when clean_directory
    void PrepareDirectory() {
        namespace fs = std::filesystem;
        fs::path d(settings_.io.directory);
        if (!fs::exists(d))
            fs::create_directories(d);
        else if (settings_.io.clean_directory)
            for (const auto& e : fs::directory_iterator(d)) fs::remove_all(e.path());
    }

    bool ShouldWrite(double time, bool force = false) const {
        if (force) return true;
        const double tol = 1e-9 * dt_snap_;
        return time >= next_snap_ - tol;
    }

    // "loose" never terminates in the reference (its loop condition is always true); it is
    // treated here as "snap", like every other value
    void AdvanceSnapshot() {
        next_snap_ += dt_snap_;
        ++idx_;
    }

    int SnapIdx() const { return idx_; }

    // X, A: AoS in the mesh's solver order
    bool WriteVTK(const zdm_mesh* m, const std::vector<double>& X, const std::vector<double>& A, int nq, int na,
                  double time, std::string* err) {
        char base[512];
        std::snprintf(base, sizeof(base), "%s-%03d.vtu", settings_.io.filename.c_str(), idx_);
        const std::string path = settings_.io.directory + "/" + base;
        if (!write_vtu(path, m, X, A, nq, na)) {
            if (err) *err = "cannot write " + path;
            return false;
        }
        log_.push_back({base, time});
        return write_series(err);
    }

private:
    struct Snap {
        std::string name;
        double time;
    };
    const Settings& settings_;
    double dt_snap_ = 0, next_snap_ = 0;
    int idx_ = 0;
    std::vector<Snap> log_;

    bool write_series(std::string* err) {
        const std::string path = settings_.io.directory + "/" + settings_.io.filename + ".vtu.series";
        std::ofstream f(path);
        if (!f.is_open()) {
            if (err) *err = "cannot write " + path;
            return false;
        }
        f << "{\n  \"file-series-version\" : \"1.0\",\n  \"files\" : [\n";
        for (size_t i = 0; i < log_.size(); ++i) {
            f << "    { \"name\" : \"" << log_[i].name << "\", \"time\" : " << std::scientific << std::setprecision(6)
              << log_[i].time << " }";
            if (i + 1 < log_.size()) f << ",";
            f << "\n";
        }
        f << "  ]\n}\n";
        return true;
    }

    static bool write_vtu(const std::string& path, const zdm_mesh* m, const std::vector<double>& X,
                          const std::vector<double>& A, int nq, int na) {
        const int nc = m->ncells, nv = m->nverts, mv = m->max_cell_verts;
        // cells in the mesh file's order: file cell i = solver cell inv[i]
        std::vector<int> inv(nc);
        for (int c = 0; c < nc; ++c) inv[m->cell_perm[c]] = c;
        std::vector<int64_t> conn, offs;
        std::vector<uint8_t> types;
        conn.reserve((size_t)nc * mv);
        for (int i = 0; i < nc; ++i) {
            const int c = inv[i];
            int k = 0;
            for (; k < mv && m->cell_verts[(size_t)c * mv + k] >= 0; ++k) conn.push_back(m->cell_verts[(size_t)c * mv + k]);
            offs.push_back((int64_t)conn.size());
            types.push_back(m->dim == 1 ? 3 : (k == 3 ? 5 : (k == 4 ? 9 : 7)));   // line, triangle, quad, polygon
        }
        std::vector<double> pts((size_t)nv * 3), st((size_t)nc * nq), ax((size_t)nc * (na > 0 ? na : 1), 0.0);
        for (size_t j = 0; j < pts.size(); ++j) pts[j] = m->verts[j];
        for (int i = 0; i < nc; ++i) {
            const int c = inv[i];
            for (int q = 0; q < nq; ++q) st[(size_t)i * nq + q] = X[(size_t)c * nq + q];
            for (int q = 0; q < na; ++q) ax[(size_t)i * na + q] = A[(size_t)c * na + q];
        }
        struct Block {
            const void* p;
            uint64_t bytes;
        };
        const int nab = na > 0 ? na : 1;
        std::vector<Block> blocks = {{pts.data(), pts.size() * 8},        {conn.data(), conn.size() * 8},
                                     {offs.data(), offs.size() * 8},      {types.data(), types.size()},
                                     {st.data(), st.size() * 8},          {ax.data(), (uint64_t)nc * nab * 8}};
        std::vector<uint64_t> at(blocks.size());
        uint64_t o = 0;
        for (size_t b = 0; b < blocks.size(); ++b) {
            at[b] = o;
            o += 8 + blocks[b].bytes;
        }
        std::ofstream f(path, std::ios::binary);
        if (!f.is_open()) return false;
        f << "<?xml version=\"1.0\"?>\n"
          << "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
          << "  <UnstructuredGrid>\n"
          << "    <Piece NumberOfPoints=\"" << nv << "\" NumberOfCells=\"" << nc << "\">\n"
          << "      <Points>\n"
          << "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" << at[0] << "\"/>\n"
          << "      </Points>\n"
          << "      <Cells>\n"
          << "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"appended\" offset=\"" << at[1] << "\"/>\n"
          << "        <DataArray type=\"Int64\" Name=\"offsets\" format=\"appended\" offset=\"" << at[2] << "\"/>\n"
          << "        <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" offset=\"" << at[3] << "\"/>\n"
          << "      </Cells>\n"
          << "      <CellData>\n"
          << "        <DataArray type=\"Float64\" Name=\"State\" NumberOfComponents=\"" << nq
          << "\" format=\"appended\" offset=\"" << at[4] << "\"/>\n"
          << "        <DataArray type=\"Float64\" Name=\"Aux\" NumberOfComponents=\"" << nab
          << "\" format=\"appended\" offset=\"" << at[5] << "\"/>\n"
          << "      </CellData>\n"
          << "    </Piece>\n"
          << "  </UnstructuredGrid>\n"
          << "  <AppendedData encoding=\"raw\">\n   _";
        for (const Block& b : blocks) {
            f.write(reinterpret_cast<const char*>(&b.bytes), 8);
            f.write(reinterpret_cast<const char*>(b.p), (std::streamsize)b.bytes);
        }
        f << "\n  </AppendedData>\n</VTKFile>\n";
        return (bool)f;
    }
};

}  // namespace zoomy
