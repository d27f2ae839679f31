"""C ABI checks that need no GPU: the library loads, exports every function
include/nbody.h declares, and rejects invalid arguments before touching a
device."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from nbody2d import abi

REPO = Path(__file__).resolve().parents[1]


def declared_functions():
    txt = (REPO / "include" / "nbody.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(nb_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(abi.EXPORTS)


def test_struct_layouts_match_header():
    assert C.sizeof(abi.Params) == 4 * 8 + 16 * 4  # 15 int32 fields + padding to 8
    assert C.sizeof(abi.Info) == 11 * 8
    assert C.sizeof(abi.Tuning) == 4 * 4 + 2 * 8 + 4 * 4
    assert abi.NODE_DTYPE.itemsize == 4 * 4 + 4 * 8 + 4 * 4 + 4 * 4


def test_params_default():
    p = abi.params()
    assert (p.G, p.dt, p.theta, p.softening) == (1.0, 1e-3, 0.5, 0.01)
    assert p.algorithm == abi.ALGO_BARNES_HUT and p.integrator == abi.INTEG_KD
    assert p.use_graph == 1


@pytest.mark.parametrize("field,value", [("dt", 0.0), ("dt", float("nan")), ("theta", -1.0),
                                         ("softening", -0.1), ("G", float("inf")),
                                         ("algorithm", 7), ("integrator", 3), ("far_field", 9),
                                         ("force_kernel", 2), ("mac", -1), ("cells", 5),
                                         ("walk_stack", 512), ("walk_bodies", 16),
                                         ("list_flush", 32), ("list_flush", 256)])
def test_create_rejects_invalid_params(field, value):
    m = np.ones(4)
    x = np.zeros((4, 2))
    with pytest.raises(abi.NBodyError) as e:
        abi.Context(m, x, x, **{field: value})
    assert e.value.status == abi.NB_ERR_INVALID_ARG


@pytest.mark.parametrize("kw", [dict(traversal=abi.TRAV_EXACT),
                                dict(algorithm=abi.ALGO_ALL_PAIRS, traversal=abi.TRAV_GROUP)])
def test_learned_far_field_needs_group_barnes_hut(kw):
    m = np.ones(4)
    x = np.zeros((4, 2))
    with pytest.raises(abi.NBodyError) as e:
        abi.Context(m, x, x, far_field=abi.FAR_LEARNED, **kw)
    assert e.value.status == abi.NB_ERR_NOT_SUPPORTED


@pytest.mark.parametrize("kw", [
    dict(cells=abi.CELLS_REF_CPU),  # rectangular cells need the diagonal test
    dict(far_field=abi.FAR_LEARNED, traversal=abi.TRAV_GROUP, force_kernel=abi.FORCE_REF_CPU),
    dict(far_field=abi.FAR_LEARNED, traversal=abi.TRAV_GROUP, mac=abi.MAC_DIAG),
])
def test_unsupported_physics_combinations(kw):
    m = np.ones(4)
    x = np.zeros((4, 2))
    with pytest.raises(abi.NBodyError) as e:
        abi.Context(m, x, x, **kw)
    assert e.value.status == abi.NB_ERR_NOT_SUPPORTED


def test_reference_cells_hold_one_system():
    m = np.ones(4)
    x = np.zeros((4, 2))
    with pytest.raises(abi.NBodyError) as e:
        abi.Context(m, x, x, offsets=[0, 2, 4], traversal=abi.TRAV_GROUP,
                    cells=abi.CELLS_REF_CPU, mac=abi.MAC_DIAG)
    assert e.value.status == abi.NB_ERR_NOT_SUPPORTED


def test_null_arrays_and_null_ctx_are_rejected():
    L = abi.lib()
    p = abi.params()
    h = C.c_void_p()
    assert L.nb_create(C.byref(p), 5, None, None, None, C.byref(h)) == abi.NB_ERR_INVALID_ARG
    assert L.nb_step(None, 1) == abi.NB_ERR_INVALID_ARG
    assert L.nb_get_state(None, None, None) == abi.NB_ERR_INVALID_ARG
    assert L.nb_status_string(abi.NB_ERR_OVERFLOW) == b"traversal stack overflow"


@pytest.mark.parametrize("offsets,kw,status", [
    ([0, 2, 4], dict(traversal=abi.TRAV_EXACT), abi.NB_ERR_NOT_SUPPORTED),
    ([0, 2, 4], dict(algorithm=abi.ALGO_ALL_PAIRS, traversal=abi.TRAV_GROUP),
     abi.NB_ERR_NOT_SUPPORTED),
    ([0, 2, 2, 4], dict(traversal=abi.TRAV_GROUP), abi.NB_ERR_INVALID_ARG),
    ([1, 2, 4], dict(traversal=abi.TRAV_GROUP), abi.NB_ERR_INVALID_ARG),
])
def test_batch_rejects_bad_offsets_and_unsupported_modes(offsets, kw, status):
    m = np.ones(4)
    x = np.zeros((4, 2))
    with pytest.raises(abi.NBodyError) as e:
        abi.Context(m, x, x, offsets=offsets, **kw)
    assert e.value.status == status


def test_batch_offsets_must_cover_the_arrays():
    m = np.ones(4)
    x = np.zeros((4, 2))
    with pytest.raises(ValueError):
        abi.Context(m, x, x, offsets=[0, 3, 5], traversal=abi.TRAV_GROUP)
