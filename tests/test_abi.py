"""The C-ABI library loads on the CPU host and exports every symbol declared
in include/hdbscan_mi355x.h (no GPU calls here)."""
import ctypes
import re
from pathlib import Path

from gpu_hdbscan_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "hdbscan_mi355x.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hdb_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for must in ("hdb_nndescent", "hdb_knn_merge", "hdb_build_knng_blocked", "hdb_hub_counts", "hdb_antihubs",
                 "hdb_core_distances", "hdb_mrd_graph", "hdb_mst", "hdb_sort_edges", "hdb_pipeline", "hdb_labels"):
        assert must in syms


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_prototype_table_covers_header():
    assert sorted(_lib.PROTOTYPES) == declared_symbols()


def test_struct_layouts():
    assert _lib.NND_ELEMENT.itemsize == 8
    assert _lib.MST_EDGE.itemsize == 12
    assert ctypes.sizeof(_lib.HdbParams) == 64
    assert ctypes.sizeof(_lib.HdbPipelineOpts) == 32


def test_params_default_are_reference_constants():
    from gpu_hdbscan_amd.api import default_params
    p = default_params()
    # gpuknn/nndescent.cuh:9-16, HDBSCAN/structs/hdbscan_elements.cuh:10-11
    assert (p.k, p.sample_num, p.nnd_iteration, p.merge_sample_num, p.merge_iteration, p.mpts) == (32, 16, 6, 12,
                                                                                                      11, 32)
