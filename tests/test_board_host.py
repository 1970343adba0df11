"""The product's board primitives (cablegen_amd/csrc/board.h, the code the
kernels run, compiled here for the host) against the oracle: prune_board
(generate.c:98-131) over its parameter range including boards of 16 equal
cells, the four moves with the lock rule under every free_formation /
ignore_f setting (board.c:151-214), and the canonical form
(board.c:314-337).  The harness is tests/native/board_vs_oracle.c."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_board_primitives_match_oracle(tmp_path):
    exe = str(tmp_path / "board_vs_oracle")
    subprocess.run(["gcc", "-O2", "-I", os.path.join(ROOT, "cablegen_amd", "csrc"), "-I",
                    os.path.join(ROOT, "oracle"), "-o", exe,
                    os.path.join(ROOT, "tests", "native", "board_vs_oracle.c"),
                    os.path.join(ROOT, "oracle", "oracle.c"), "-lpthread", "-lm"],
                   check=True, capture_output=True, text=True, timeout=120)
    r = subprocess.run([exe, "400000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok 400000 boards")
