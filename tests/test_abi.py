"""The C-ABI library loads and exports every symbol include/cospdnn.h declares (no device calls)."""
import os
import re

import cospdnn_amd as cd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "cospdnn.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cospdnn_\w+)\s*\(", text)))


def test_header_symbols_exported():
    lib = cd.lib()
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(cd.EXPORTS) == names


def test_version_and_errors():
    assert "gfx950" in cd.version()
    lib = cd.lib()
    assert lib.cospdnn_strerror(0) == b"ok"
    assert lib.cospdnn_strerror(5) == b"values not exact in FP16"


def test_host_helpers_reject_bad_arguments():
    import numpy as np
    import pytest

    with pytest.raises(cd.CospdnnError):
        cd.generate_layer(1, cd.GEN_BANDED, 100, 32, 0)  # banded needs K | N
    with pytest.raises(cd.CospdnnError):
        cd.layer_stats(4, np.array([0, 1, 2, 3, 4], np.int32), np.array([0, 1, 9, 3], np.int32),
                       np.ones(4, np.float32))  # row index out of range


def test_product_library_is_in_tree():
    assert os.path.dirname(cd.LIB_PATH) == os.path.join(ROOT, "cospdnn_amd")


def test_options_struct_matches_the_header():
    """The ctypes mirror of cospdnn_opts lines up with the C struct: the defaults the library
    writes land in the right fields, the last one included."""
    o = cd.default_opts()
    assert o.precision == cd.FP16
    assert o.guard_layer == 12 and o.guard_layer32 == 20
    assert o.use_mfma == 1 and abs(o.max_tile_waste - 64.0) < 1e-6
    assert o.dump_layer == -1
    assert o.spmm_order == 1 and o.mfma_groups_per_cu == 4
    assert o.batch_concurrent == 2 and o.guard_tail == 8
    assert o.guard_concurrent == 1
    assert o.keep_holes_ppm == 0
    assert o.spmm_gather == 1
    assert o.input_bits == 1
    assert o.weight_mask == 1
    assert o.weight_budget_mb == 0
