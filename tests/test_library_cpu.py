"""The native library on the CPU: it loads, exports every symbol include/gpu_phs.h declares,
compiles and validates trees, reads the reference's .ref format, and reports call-order and
argument errors.  No GPU compute calls here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import gpu_phs_amd as G
from gpu_phs_amd import process as P
from gpu_phs_amd import workloads as W
from refio import read_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpu_phs.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


# declared with C++ linkage (transfer_to_whizard.h:1-3): exported under their mangled names
CXX_SYMBOLS = {"get_momentum_device_pointer": "_Z27get_momentum_device_pointerv",
               "get_in_momentum_device_pointer": "_Z30get_in_momentum_device_pointerv",
               "get_ok_device_pointer": "_Z21get_ok_device_pointerv"}


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(\w+)\s*\(", text, flags=re.M))
    return sorted(CXX_SYMBOLS.get(n, n) for n in names)


def test_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    assert sum(n.startswith("c_whizard_") for n in names) == 14
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if " T " in line)
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert set(CXX_SYMBOLS.values()) <= set(names)
    lib = G.lib()
    for n in names:
        assert getattr(lib, n) is not None


def test_whizard_symbols_match_reference_names():
    """The 14 entry points of whizard_gpu_phs.cpp:12-249 (names as the Fortran binding and
    Whizard's C side use them)."""
    expected = ["c_whizard_set_particle_structure", "c_whizard_set_flavors", "c_whizard_init_mappings",
                "c_whizard_fill_mapping", "c_whizard_init_tree_structures", "c_whizard_fill_tree_structure",
                "c_whizard_init_channel_ids", "c_whizard_set_threads", "c_whizard_init_gpu_phs",
                "c_whizard_finalize_gpu_phs", "c_whizard_gen_phs_from_x_gpu",
                "c_whizard_get_momentum_device_pointer", "c_whizard_reset_generator", "c_whizard_show_module"]
    assert sorted(expected) == sorted(n for n in declared_functions() if n.startswith("c_whizard_"))


@pytest.mark.parametrize("name", ["c0", "c1", "c2", "2to6", "friends4", "friends6"])
def test_process_create(name):
    proc = {"c0": W.process_c0, "c1": W.process_c1, "c2": W.process_c2, "2to6": W.process_2to6,
            "friends4": lambda: W.process_friends(4), "friends6": lambda: W.process_friends(6)}[name]()
    ph = G.ProcessHandle(proc)
    n_out, n_x, n_ch = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert G.lib().gpu_phs_process_info(ph.handle, ctypes.byref(n_out), ctypes.byref(n_x), ctypes.byref(n_ch)) == 0
    assert (n_out.value, n_x.value, n_ch.value) == (proc.n_out, proc.n_x, proc.n_channels)


@pytest.mark.parametrize("n_out", [2, 3, 4, 5, 6])
def test_every_small_tree_has_a_specialised_kernel(n_out):
    """Each ordered tree shape of 2-6 final-state particles matches its own entry of the
    compile-time shape table (phs_shape_table.h, written by gen_shapes.py), whatever the
    particle labels and maps: the generated programs agree with the tree compiler's."""
    proc = W.process_shapes(n_out)
    shapes = G.ProcessHandle(proc).shapes()
    n_shapes = len(W.ordered_trees(list(range(n_out))))
    per = proc.n_channels // n_shapes
    assert (shapes >= 0).all()
    by_tree = shapes.reshape(n_shapes, per)
    assert (by_tree == by_tree[:, :1]).all()              # labels and maps do not matter
    assert len(set(by_tree[:, 0].tolist())) == n_shapes   # distinct shapes, distinct kernels


def test_larger_trees_and_friends_run_the_generic_kernel():
    for proc in (W.process_large(8), W.process_friends(4), W.process_friends(6)):
        assert (G.ProcessHandle(proc).shapes() == -1).all()
    for proc in (W.process_c2(), W.process_2to6()):
        assert (G.ProcessHandle(proc).shapes() >= 0).all()


def _expect_tree_error(proc, fragment):
    with pytest.raises(G.GpuPhsError) as e:
        G.ProcessHandle(proc)
    assert "(-2)" in str(e.value) and fragment in str(e.value)
    assert G.lib().gpu_phs_last_error() == -2


def test_invalid_trees_are_rejected():
    p = W.process_c1()
    p.d1[0, 14] = 5          # root 15 -> (5, 12): bincodes overlap
    _expect_tree_error(p, "inconsistent daughters")
    p = W.process_c1()
    p.has_children[0, 2] = 0  # branch 3 loses its children: leaves 1, 2 missing
    _expect_tree_error(p, "not a final particle")
    p = W.process_c1()
    p.map_ids[0, 2] = P.MAP_ONSHELL   # no msq mapping in the reference either
    _expect_tree_error(p, "no msq mapping")
    p = W.process_c1()
    p.map_ids[0, 14] = 7     # unknown id on the root: no cos-theta mapping
    _expect_tree_error(p, "no cos-theta mapping")


def test_invalid_arguments():
    p = W.process_c1()
    with pytest.raises(G.GpuPhsError):
        G.ProcessHandle(p, E1=-1.0, E2=250.0)
    assert G.lib().gpu_phs_process_destroy(12345) == -5


@pytest.mark.parametrize("name", ["ttH_head.ref", "bwbw_head.ref"])
def test_load_ref_matches_python_reader(name):
    path = os.path.join(GOLDEN, name)
    proc, _ = read_ref(path)
    h = G.load_ref(path, proc.sqrts / 2, proc.sqrts / 2)
    n = proc.n_channels
    d1 = np.zeros((n, proc.n_prt), np.int32)
    d2 = np.zeros_like(d1)
    hc = np.zeros_like(d1)
    mid = np.zeros((n, proc.n_prt_out), np.int32)
    mm = np.zeros((n, proc.n_prt_out))
    mw = np.zeros_like(mm)
    flv = np.zeros(proc.n_in + proc.n_out)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert G.lib().gpu_phs_process_trees(h, ptr(d1), ptr(d2), ptr(hc), ptr(mid), ptr(mm), ptr(mw), ptr(flv)) == 0
    assert np.array_equal(d1, proc.d1) and np.array_equal(d2, proc.d2) and np.array_equal(hc, proc.has_children)
    assert np.array_equal(mid, proc.map_ids) and np.array_equal(mm, proc.map_masses)
    assert np.array_equal(mw, proc.map_widths) and np.array_equal(flv, proc.flv_masses)


def test_load_ref_errors(tmp_path):
    with pytest.raises(G.GpuPhsError):
        G.load_ref(str(tmp_path / "missing.ref"), 250.0, 250.0)
    bad = tmp_path / "bad.ref"
    bad.write_text(" n_channels: 1\n NIN: 2\n NOUT: 4\n NTreesForest: 1\n NTreesGrove: 1\n nx: 7\n")
    with pytest.raises(G.GpuPhsError, match="nx"):
        G.load_ref(str(bad), 250.0, 250.0)


def _i(v):
    return ctypes.byref(ctypes.c_int(v))


def test_whizard_call_order_errors():
    L = G.lib()
    L.c_whizard_finalize_gpu_phs()
    L.c_whizard_init_mappings()
    assert L.gpu_phs_last_error() == -3           # before set_particle_structure
    L.c_whizard_set_particle_structure(_i(3), _i(3), _i(1), _i(8), _i(2), _i(4))
    assert L.gpu_phs_last_error() == 0
    L.c_whizard_init_gpu_phs()
    assert L.gpu_phs_last_error() == -3           # mappings and trees not set
    L.c_whizard_set_particle_structure(_i(3), _i(3), _i(1), _i(7), _i(2), _i(4))
    assert L.gpu_phs_last_error() == -1           # n_x != 3 n_out - 4
    L.c_whizard_finalize_gpu_phs()


def test_whizard_init_channel_ids_skips_interior_empty_channels():
    """whizard_gpu_phs.cpp:89-98 assigns events after an interior empty channel to that empty
    channel; here they go to the next non-empty one."""
    L = G.lib()
    limits = (ctypes.c_int * 4)(0, 3, 3, 5)      # channel 0 empty, 1: events 0-2, 2 empty, 3: 3-4
    L.c_whizard_init_channel_ids(_i(5), _i(4), limits)
    assert L.gpu_phs_last_error() == 0
    short = (ctypes.c_int * 2)(1, 2)
    L.c_whizard_init_channel_ids(_i(5), _i(2), short)
    assert L.gpu_phs_last_error() == -1


def test_whizard_structure_and_show_module(capfd):
    proc = W.process_c1()
    L = G.lib()
    L.c_whizard_set_particle_structure(_i(3), _i(3), _i(1), _i(8), _i(2), _i(4))
    flv = np.ascontiguousarray(proc.flv_masses)
    wid = np.zeros_like(flv)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    L.c_whizard_set_flavors(ptr(flv), ptr(wid))
    L.c_whizard_init_mappings()
    L.c_whizard_init_tree_structures()
    for c in range(3):
        L.c_whizard_fill_mapping(_i(c + 1), ptr(np.ascontiguousarray(proc.map_ids[c])),
                                 ptr(np.ascontiguousarray(proc.map_masses[c])),
                                 ptr(np.ascontiguousarray(proc.map_widths[c])))
        z = np.zeros(proc.n_prt, np.int32)
        L.c_whizard_fill_tree_structure(_i(c + 1), ptr(np.ascontiguousarray(proc.d1[c])),
                                        ptr(np.ascontiguousarray(proc.d2[c])),
                                        ptr(np.ascontiguousarray(proc.has_children[c])), ptr(z))
        assert L.gpu_phs_last_error() == 0
    L.c_whizard_show_module()
    out = capfd.readouterr().out
    assert "n_channels: 3" in out and "15->(3,12)" in out and "END OF MODULE" in out
    L.c_whizard_fill_mapping(_i(4), ptr(np.zeros(15, np.int32)), ptr(np.zeros(15)), ptr(np.zeros(15)))
    assert L.gpu_phs_last_error() == -1         # channel out of range
    L.c_whizard_finalize_gpu_phs()
