// Times meso's own gambit reader (MESO::fvmMesh::load_gambit: parse, faces,
// geometry and least-squares set-up, src/mesh/mesh.cpp:4-11 and
// src/mesh/mesh_reader.cpp:8-139), compiled from the reference sources by
// oracle/Makefile.ref into oracle/_ref/meso_mesh_timer.
//
// TEST INFRASTRUCTURE ONLY: the front-end speed test compares the library's
// case reader against it. Nothing in the product uses it.
//
// usage: meso_mesh_timer <mesh.neu>   -> prints "cells <n> seconds <t>"
#include "mesh/mesh.h"

#include <chrono>
#include <cstdio>

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: meso_mesh_timer <mesh.neu>\n");
        return 2;
    }
    MESO::MPI::Initialize(&argc, &argv);
    const auto t0 = std::chrono::steady_clock::now();
    MESO::fvmMesh::Mesh mesh = MESO::fvmMesh::load_gambit(argv[1], 1.0);
    const auto t1 = std::chrono::steady_clock::now();
    std::printf("cells %d seconds %.6f\n", mesh.NCELL, std::chrono::duration<double>(t1 - t0).count());
    MESO::MPI::Finalize();
    return 0;
}
