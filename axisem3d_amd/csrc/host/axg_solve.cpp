// axg_solve.cpp -- a C++ caller of the Domain / Newmark facade (axg_domain.hpp),
// the way the reference's axisem_main drives its solver (SOLVER/src/axisem.cpp:
// 170-204): build the mesh, release it into a Domain, Newmark::solve, and
// report exceptions in the "Class::method || message" form.
//
//   axg_solve <workload 0-3> [nstep] [nelem_theta] [seismogram output file]
//
// The output file holds the float32 seismograms [nstep][nrec][3].
#include <cstdio>
#include <cstdlib>
#include <exception>

#include "axg_domain.hpp"

int main(int argc, char **argv) {
    int which = argc > 1 ? std::atoi(argv[1]) : 0;
    int nstep = argc > 2 ? std::atoi(argv[2]) : 0;
    int ntheta = argc > 3 ? std::atoi(argv[3]) : 0;
    const char *out = argc > 4 ? argv[4] : nullptr;
    axg_mesh *mesh = nullptr;
    try {
        axg_shell_config cfg;
        axg::throwIfFailed(axg_shell_config_default(which, &cfg), "axg_solve");
        if (nstep > 0) cfg.nstep = nstep;
        if (ntheta > 0) cfg.nelem_theta = ntheta;
        axg::throwIfFailed(axg_mesh_build_shell(&cfg, &mesh), "Mesh::release");
        const axg_mesh_desc *desc = axg_mesh_get_desc(mesh);
        axg::Domain *domain = new axg::Domain(desc);
        axg::Newmark newmark(domain, 100, 100);
        newmark.solve(1);
        std::vector<float> seis = domain->seismograms();
        delete domain;
        if (out) {
            FILE *f = std::fopen(out, "wb");
            if (!f || std::fwrite(seis.data(), sizeof(float), seis.size(), f) != seis.size())
                throw std::runtime_error("axg_solve || cannot write the seismograms.");
            std::fclose(f);
        }
        std::printf("axg_solve: %d steps, %d stations\n", desc->nstep, desc->nrec);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        if (mesh) axg_mesh_free(mesh);
        return 1;
    }
    axg_mesh_free(mesh);
    return 0;
}
