"""The reference-side binding shown in the integration guide compiles against the
reference's own headers (src/BTEMain.cpp's includes) and include/giftbte.h.

Nothing of the reference is copied: the test writes a translation unit that
includes the reference's headers where they lie and holds only the binding
text from that guide plus a call of it, then runs g++ -fsyntax-only with
the flags oracle/ref.mk compiles the reference with."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
MPI_INC = "/opt/conda/include"


def _binding_text():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"// --- giftbte binding:.*?// --- end of giftbte binding ---", doc, re.S)
    assert m, "binding block missing from INTEGRATION.md"
    return m.group(0)


def test_binding_block_present():
    text = _binding_text()
    assert "gbte_fourier_solve" in text and "gbte_solve(" in text and "gbte_get_fields" in text


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "includes")) or shutil.which("g++") is None
                    or not os.path.exists(os.path.join(MPI_INC, "mpi.h")),
                    reason="reference headers, g++ or mpi.h absent")
def test_binding_compiles_against_reference_headers(tmp_path):
    src = tmp_path / "binding_check.cpp"
    src.write_text(
        "// the includes of src/BTEMain.cpp\n"
        "#include <iostream>\n#include <fstream>\n#include <sstream>\n#include <string>\n"
        '#include "StaticBTESynthetic/StaticBTESynthetic.h"\n#include "StaticFourier/StaticFourier.h"\n'
        '#include "BTEMesh/Distribute.h"\n#include "Solution/Solution.h"\n'
        "using namespace std;\n#define USE_GIFTBTE_HIP 1\n"
        + _binding_text() +
        "\nint binding_call(SolutionAll &sa, DistributeMesh *dm, string State) {\n"
        "    int world_rank = 0, Num_Max_Iter = 10;\n"
        "    double error_temp_limit = 1e-8, error_flux_limit = 1e-4;\n"
        "#ifdef USE_GIFTBTE_HIP\n"
        "    if (State != \"Transient\" && giftbte_steady(sa, dm, world_rank, Num_Max_Iter,\n"
        "                                               error_temp_limit, error_flux_limit)) {\n"
        "    } else\n#endif\n"
        "    if (State != \"Transient\") return 1;\n"
        "    return 0;\n}\n")
    cmd = ["g++", "-std=c++14", "-fsyntax-only", "-w", f"-I{REF}/includes", f"-I{MPI_INC}",
           f"-I{os.path.join(ROOT, 'include')}", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
