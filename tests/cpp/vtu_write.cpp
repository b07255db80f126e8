// Writes two VTK snapshots of a mesh through zoomy::IOManager with synthetic cell data
// (State[c][q] = 10 * input_cell + q, Aux[c] = -input_cell) for tests/test_host_cpp.py.
//   vtu_write <mesh.msh> <out_dir>
#include <iostream>
#include <vector>

#include "zoomy/io.hpp"

int main(int argc, char** argv) {
    if (argc < 3) return 1;
    zdm_mesh* m = nullptr;
    if (zdm_mesh_read_msh(argv[1], 1, &m)) {
        std::cerr << zdm_mesh_last_error() << std::endl;
        return 2;
    }
    zoomy::Settings s;
    s.io.directory = argv[2];
    s.io.filename = "snap";
    s.io.snapshots = 2;
    s.solver.t_end = 1.0;
    zoomy::IOManager io(s);
    io.PrepareDirectory();
    const int nq = 3, na = 1, nc = m->ncells;
    std::vector<double> X((size_t)nc * nq), A(nc);
    for (int c = 0; c < nc; ++c) {
        const int i = m->cell_perm[c];
        for (int q = 0; q < nq; ++q) X[(size_t)c * nq + q] = 10.0 * i + q;
        A[c] = -i;
    }
    std::string err;
    for (double t : {0.0, 0.3, 0.5, 0.9, 1.0}) {
        if (!io.ShouldWrite(t)) continue;
        if (!io.WriteVTK(m, X, A, nq, na, t, &err)) {
            std::cerr << err << std::endl;
            return 3;
        }
        io.AdvanceSnapshot();
    }
    zdm_mesh_destroy(m);
    return 0;
}
