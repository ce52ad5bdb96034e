"""The reference-side binding compiles and links against meso itself.

meso-lts_amd/integration/meso_hip_solver.h is the solver class a maintainer
adds to meso (BasicSolver contract, include/solver/solver.h:22-46; time loop of
src/meso/exec.cpp:141-165; dispatch of src/meso/main.cpp:47-57). This test
compiles it with meso's own headers, links it with meso's own objects (the ones
oracle/Makefile.ref builds from the reference sources) and libmeso_hip.so, and
runs the program on a case: without a GPU the class reaches meso_hip_create,
which reports that no device is there, and the error comes back through meso's
logger and a non-zero exit. Skipped where the reference is absent (it is never
on a GPU box).
"""
import glob
import os
import subprocess

import pytest

from conftest import gpu_available
from meso_hip import capi
from meso_hip import meshgen as mg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
REF_OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")
MPI_INC = "/opt/conda/include"
MPI_LIB = os.path.join(ROOT, "oracle", "_ref", "lib", "libmpi.so.12")

pytestmark = pytest.mark.skipif(
    not (os.path.isdir(os.path.join(REF, "include")) and os.path.isdir(REF_OBJ) and os.path.exists(MPI_LIB)),
    reason="needs the reference sources and oracle/_ref (make -f oracle/Makefile.ref)")

# the solver dispatch of src/meso/main.cpp:47-57 with the added solver names
MAIN = r'''
#include "meso.h"
#include "meso_hip_solver.h"

int MESO::Solver::solver_state = 0;  // defined by src/meso/exec.cpp, which this link leaves out

int main(int argc, char **argv) {
    MESO::ArgParser parser(argc, argv);
    auto case_file = parser.parse_param<MESO::String>("case", "<case-file>", false);
    MESO::MPI::Initialize(&argc, &argv);
    int status = 0;
    try {
        MESO::Solver::Config config(case_file);
        auto solver = config.get<std::string>("solver", "<solver>", false);
        if (solver == "cdugks@shakhov-hip" || solver == "cdugks@incompressible-hip" ||
            solver == "dugks@incompressible-hip")
            status = MESO::Solver::handle_solver<MESO::Solver::MESO_HIP_SOLVER>(parser, config);
        else
            logger.warn << "Load solver: " << solver << " failed." << std::endl;
    } catch (const std::exception &e) {
        std::cout << "solver error: " << e.what() << std::endl;
        status = 1;
    }
    MESO::MPI::Finalize();
    return status;
}
'''

FLAGS = ["-std=c++17", "-O1", "-DMPICH_SKIP_MPICXX", "-DOMPI_SKIP_MPICXX", "-I" + os.path.join(REF, "include"),
         "-I" + MPI_INC, "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "meso-lts_amd", "integration")]


def test_solver_class_compiles_against_meso(tmp_path):
    src = tmp_path / "main.cpp"
    src.write_text(MAIN)
    r = subprocess.run(["g++", *FLAGS, "-Wall", "-Wno-unused-variable", "-fsyntax-only", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_solver_class_links_and_runs_meso_loop(tmp_path):
    src = tmp_path / "main.cpp"
    src.write_text(MAIN)
    obj = tmp_path / "main.o"
    r = subprocess.run(["g++", *FLAGS, "-w", "-c", str(src), "-o", str(obj)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    # meso's own objects, not the drivers of oracle/ (ref_*.cpp, each with a main)
    objs = [o for o in glob.glob(os.path.join(REF_OBJ, "**", "*.o"), recursive=True)
            if not os.path.basename(o).startswith("ref_")]
    exe = tmp_path / "meso_hip_solver"
    libdir = os.path.dirname(capi.LIB_PATH)
    r = subprocess.run(["g++", "-o", str(exe), str(obj), *objs, MPI_LIB, "-L" + libdir, "-lmeso_hip",
                        "-Wl,-rpath," + libdir + ":" + os.path.dirname(MPI_LIB)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    # a small Shakhov cavity under the new solver name
    case = mg.shakhov_cavity_case(str(tmp_path / "case"), mg.quad_cavity(4, jitter=0.1, seed=1), ("nc", 1, 7, 4.5),
                                  mg.viscosity_from_kn(0.075), 0.2, max_step=3, name="hipcase")
    text = open(case).read().replace("cdugks@shakhov", "cdugks@shakhov-hip")
    open(case, "w").write(text)
    r = subprocess.run([str(exe), "--case", "case.txt"], cwd=os.path.dirname(case), capture_output=True, text=True,
                       timeout=120)
    if gpu_available():
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert os.path.exists(os.path.join(os.path.dirname(case), "hipcase", "latest"))
    else:
        # the class got as far as the device context and reported its error
        assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
        assert "no HIP device" in r.stdout, r.stdout[-2000:]
