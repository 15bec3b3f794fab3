"""The product builds from the committed sources alone.

- every public header compiles on its own as C99 and as C++17;
- the C client of the drop-in layer compiles against those headers;
- a clean build of the library into an empty directory compiles every source
  file afresh, links, and exports every function the headers declare.
No GPU is needed: hipcc cross-compiles for gfx950.
"""
import glob
import os
import subprocess
import time

import pytest

from ifos_amd import build as B
from test_library import declared_functions

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def _ok(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout
    return r.stdout


def test_headers_found():
    assert len(HEADERS) >= 5


@pytest.mark.parametrize("header", [os.path.basename(h) for h in HEADERS])
@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_compiles_standalone(header, lang):
    path = os.path.join(ROOT, "include", header)
    if lang == "c":
        _ok(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c", path])
    else:
        _ok(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-x", "c++", path])


def test_header_included_twice_and_together(tmp_path):
    src = tmp_path / "all.c"
    src.write_text("".join(f'#include "{os.path.basename(h)}"\n' for h in HEADERS * 2) + "int main(void){return 0;}\n")
    _ok(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"), str(src)])


MPI_INCLUDE = "/opt/conda/include"


@pytest.mark.skipif(not os.path.exists(os.path.join(MPI_INCLUDE, "mpi.h")), reason="no mpi.h in the image")
def test_dropin_header_beside_the_reference_prototypes(tmp_path):
    """A driver that includes mpi.h, the reference's prototypes of
    exchange_v / exchange_s (src/fd.h:116-124, restated here) and the drop-in
    header compiles: the declarations agree."""
    src = tmp_path / "driver.c"
    src.write_text(
        "#include <mpi.h>\n"
        "void exchange_v(float ** vx, float ** vy, float ** vz, float ** bufferlef_to_rig,\n"
        "    float ** bufferrig_to_lef, float ** buffertop_to_bot, float ** bufferbot_to_top,\n"
        "    MPI_Request * req_send, MPI_Request * req_rec, int wavetyp_start);\n"
        "void exchange_s(float ** sxx, float ** syy, float ** sxy, float ** sxz, float ** syz,\n"
        "    float ** bufferlef_to_rig, float ** bufferrig_to_lef, float ** buffertop_to_bot,\n"
        "    float ** bufferbot_to_top, MPI_Request * req_send, MPI_Request * req_rec, int wavetyp_start);\n"
        '#include "ifos2d_dropin.h"\n'
        "int main(void){return 0;}\n")
    _ok(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I" + MPI_INCLUDE,
         "-I" + os.path.join(ROOT, "include"), str(src)])


@pytest.mark.parametrize("client", ["dropin_forward.c", "engine_gradient.c"])
def test_c_client_compiles(client):
    _ok(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"),
         "-I" + os.path.join(ROOT, "ifos_amd", "csrc"), os.path.join(ROOT, "tests", "c", client)])


def test_clean_build_rebuilds_every_object(tmp_path):
    t0 = time.time()
    obj = tmp_path / "obj"
    lib = tmp_path / "lib" / "libifos2d_hip.so"
    B.build(obj_dir=str(obj), lib_path=str(lib))
    objs = sorted(obj.glob("*.o"))
    want = sorted(os.path.splitext(os.path.basename(s))[0] + ".o" for s in B.sources())
    assert [o.name for o in objs] == want
    for o in objs:
        assert o.stat().st_mtime >= t0 - 1, f"{o.name} not rebuilt"
    assert lib.stat().st_mtime >= t0 - 1
    syms = set(l.split()[-1] for l in _ok(["nm", "-D", "--defined-only", str(lib)]).splitlines() if l.strip())
    missing = [n for n in declared_functions() if n not in syms]
    assert not missing, f"declared but not exported: {missing}"
