"""bench.py's reading of the committed main-kernel traffic: reported only
while the kernel's sources hash to what the counters were collected on."""
import os
import shutil

import bench


def _copy_tree(tmp_path):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for rel in ("profiles/gemm_packed_bf16x3_8192_traffic.txt", "mojoblas_amd/csrc/gemm_packed.hpp",
                "mojoblas_amd/csrc/common.hpp"):
        dst = tmp_path / rel
        dst.parent.mkdir(parents=True, exist_ok=True)
        shutil.copy(os.path.join(root, rel), dst)
    return tmp_path


def test_committed_traffic_matches_current_sources():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    traffic, note = bench.committed_traffic(root)
    assert traffic is not None and note is None
    assert 8.05e8 < traffic < 1e10  # above the algorithmic bytes, below any plausible re-read factor


def test_changed_kernel_source_makes_traffic_null(tmp_path):
    root = _copy_tree(tmp_path)
    assert bench.committed_traffic(str(root))[0] is not None
    with open(root / "mojoblas_amd/csrc/gemm_packed.hpp", "a") as f:
        f.write("\n// changed\n")
    traffic, note = bench.committed_traffic(str(root))
    assert traffic is None and note.startswith("stale")


def test_missing_profile_gives_none(tmp_path):
    assert bench.committed_traffic(str(tmp_path)) == (None, None)
