"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: it is the
arbiter of every bit-exact test and indexes raw arrays everywhere (the coarse
operator's position and marker arrays, the cellCells positions), so its own
memory safety is checked: oracle/build.sh sanitize builds oracle/sanitize_main
from the oracle sources with -fsanitize=address,undefined (no recovery), which
runs the KAT-226 solve with every cycle type and smoother, and the assembly,
gradient and solve on hexahedral, 2-D quadrilateral and tetrahedral meshes."""
import os
import subprocess

import numpy as np
import pytest

import fluent_golden as FG
import hexmesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def sanitized(tmp_path_factory):
    try:
        subprocess.run(["sh", os.path.join(ROOT, "oracle", "build.sh"), "sanitize"], check=True,
                       capture_output=True, timeout=300)
    except (subprocess.CalledProcessError, FileNotFoundError) as e:
        pytest.skip("no sanitizer toolchain: %s" % e)
    return os.path.join(ROOT, "oracle", "_sanitize", "sanitize_main")


def _write_mesh(path, m):
    nodes = np.ascontiguousarray(m["nodes"], np.float64).reshape(-1)
    nf = m["face_cells"].size // 2
    fnptr = np.ascontiguousarray(m.get("face_nptr", np.arange(nf + 1) * 4), np.int32)
    fnodes = np.ascontiguousarray(m["face_nodes"], np.int32)
    gp = np.ascontiguousarray(m["group_ptr"], np.int32)
    hdr = np.array([m.get("dim", 3), nodes.size // 3, nf, m["n_cells"], m["n_ghost"], gp.size - 1, fnodes.size],
                   np.int32)
    with open(path, "wb") as f:
        for a in (hdr, nodes, fnptr, fnodes, np.ascontiguousarray(m["face_cells"], np.int32), gp):
            f.write(a.tobytes())


@pytest.mark.parametrize("case", ["hex", "cav32.cas", "cav_tetra.cas"])
def test_oracle_runs_clean_under_sanitizers(sanitized, tmp_path, case):
    if case == "hex":
        m = hexmesh.hex_mesh(4, 0.2, 3)
    else:
        m = FG.read_mesh(case).arrays()
    mf = str(tmp_path / "mesh.bin")
    _write_mesh(mf, m)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([sanitized, os.path.join(GOLD, "MatrixMarket226.dat"), os.path.join(GOLD, "rhs226.dat"), mf],
                       capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout
