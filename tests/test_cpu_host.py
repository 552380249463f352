"""CPU-only checks of the library's host side: exported C ABI, scene builders,
cameras and argument validation (no GPU calls)."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import litert_amd.api as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return set(re.findall(r"^\s*(?:int|const char\*|void)\s+(lrt_\w+)\s*\(", text, re.M))


def test_library_exports_every_declared_symbol():
    declared = _declared("lrt.h")
    assert len(declared) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], stdout=subprocess.PIPE, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = declared - exported
    assert not missing, missing
    for name in declared:  # and ctypes can resolve each one
        assert hasattr(L.lib, name)


def test_struct_layouts_match_reference_sizes():
    # sdf_scene.h:136-161, render_settings.h:86-95, cbvh.h:21-33
    assert C.sizeof(L.SdfSBSHeader) == 16
    assert C.sizeof(L.MultiRenderPreset) == 28
    assert C.sizeof(L.CRT_Hit) == 32


def test_sbs_builder_keeps_reference_border_rule():
    s = L.displaced_icosphere(3, 0.7, 0.05, 2, seed=5, lod=16, brick_size=4, brick_pad=1, bytes_per_value=4)
    assert s.header.aux_data == L.SDF_SBS_NODE_LAYOUT_DX
    v = 4 + 2 + 1
    words = v ** 3
    assert np.array_equal(s.nodes[:, 2], words * np.arange(s.nodes.shape[0]))
    d_max = 2 * np.sqrt(3) / 16
    q = s.values.reshape(-1, words).astype(np.float64)
    vals = -d_max + (2 * d_max / 4294967295.0) * q
    # sparse_octree_builder.cpp:49-53: min <= 0 and max > -d_max
    assert np.all(vals.min(axis=1) <= 1e-9)
    assert np.all(vals.max(axis=1) > -d_max)
    # and every cell crossing the zero level set was found: bricks along a radius
    pos = np.stack([s.nodes[:, 0] >> 16, s.nodes[:, 0] & 0xFFFF, s.nodes[:, 1] >> 16], 1)
    centre = -1 + (pos + 0.5) * (2.0 / 16)
    r = np.linalg.norm(centre, axis=1)
    band = d_max + np.sqrt(3) * (2.0 / 16)  # keep band + brick (with pad) half-diagonals
    assert r.min() > 0.7 - 0.05 - band and r.max() < 0.7 + 0.05 + band


def test_voxel_diagonal_quantisation_flag():
    s = L.displaced_icosphere(3, 0.7, 0.05, 2, seed=5, lod=16, brick_size=4, brick_pad=0, bytes_per_value=1, voxel_diag_scale=True)
    assert s.header.aux_data == (L.SDF_SBS_NODE_LAYOUT_DX | 1)
    assert s.values.size == s.nodes.shape[0] * 32  # 125 bytes -> 32 words per brick


def test_grid_sbs_layout():
    s = L.grid_sbs("torus", 0.6, 0.2, 0.0, 1, seed=1, brick_count=3, brick_size=2, with_neighbors=True)
    assert s.header.aux_data == L.SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN
    per = 27 + 8 + 27
    assert s.values.size == 27 * per
    p_count = 3 * 2 + 1
    assert s.values_f.size == p_count ** 3 + 3 * 4 ** 3
    # distance index of brick (1,1,1) corner (0,0,0) -> lattice point (2,2,2)
    b = (1 * 3 + 1) * 3 + 1
    assert s.values[b * per] == (2 * p_count + 2) * p_count + 2
    # neighbour table: slot 9*(dx+1)+3*(dy+1)+(dz+1)
    nb = s.values[b * per + 27 + 8: b * per + 27 + 8 + 27]
    assert nb[0] == 0 and nb[26] == 26 and nb[13] == b
    corner = s.values[27 + 8: 27 + 8 + 27]  # brick (0,0,0): negative offsets are missing
    assert corner[0] == L.INVALID_IDX and corner[26] == (1 * 3 + 1) * 3 + 1


def test_orbit_cameras_look_at_origin():
    for i in range(8):
        wv, pr = L.orbit_camera(2, i)
        M = wv.reshape(4, 4).T  # column-major -> row-major
        eye = -M[:3, :3].T @ M[:3, 3]
        assert abs(np.linalg.norm(eye) - 3.0) < 1e-4
        fwd = -M[2, :3]  # camera looks down -z
        assert np.dot(fwd, -eye / np.linalg.norm(eye)) > 1 - 1e-5


def test_invalid_arguments_are_rejected():
    h = C.c_void_p()
    assert L.lib.lrt_create(L.DEVICE_CPU, 0, C.byref(h)) != 0  # CPU device is not served by this library
    assert L.lib.lrt_sbs_grid(1, 0.6, 0.2, 0.0, 1, 1, 0, 4, 0, C.byref(h)) != 0
    assert b"bad" in L.lib.lrt_last_error()
    assert L.lib.lrt_sbs_load_file(b"/nonexistent/file.bin", C.byref(h)) != 0


def test_displaced_torus_builder():
    s = L.displaced_torus(0.6, 0.2, 0.02, 2, seed=3, lod=16, brick_size=4, brick_pad=0, bytes_per_value=1)
    assert s.header.aux_data == L.SDF_SBS_NODE_LAYOUT_DX and s.nodes.shape[0] > 50
    # brick centres lie near the torus surface: |dist to the tube centre circle - r| within the keep band
    pos = np.stack([s.nodes[:, 0] >> 16, s.nodes[:, 0] & 0xFFFF, s.nodes[:, 1] >> 16], 1)
    c = -1 + (pos + 0.5) * (2.0 / 16)
    q = np.sqrt(np.hypot(c[:, 0], c[:, 1]) ** 2) - 0.6
    d = np.abs(np.hypot(q, c[:, 2]) - 0.2)
    assert d.max() < 0.02 + 2 * np.sqrt(3) / 16 + np.sqrt(3) / 16


def test_mesh_builder_matches_icosphere_builder():
    """lrt_sbs_from_mesh on the displaced icosphere's own mesh gives the scene
    lrt_sbs_displaced_icosphere builds (same distance function and rule)."""
    v, t = L.mesh_displaced_icosphere(4, 0.8, 0.05, 3, 11)
    assert v.shape == (2562, 3) and t.shape == (5120, 3) and t.max() < v.shape[0]
    a, ms = L.sbs_from_mesh(v, t, lod=32, brick_size=4, brick_pad=1, bytes_per_value=4)
    b = L.displaced_icosphere(4, 0.8, 0.05, 3, seed=11, lod=32, brick_size=4, brick_pad=1, bytes_per_value=4)
    assert ms > 0 and a.nodes.shape[0] > 500
    assert np.array_equal(a.nodes, b.nodes) and np.array_equal(a.values, b.values)
    with pytest.raises(L.LrtError):
        L.sbs_from_mesh(v, t, lod=48)  # not a power of two
    bad = t.copy()
    bad[0, 0] = v.shape[0]
    with pytest.raises(L.LrtError):
        L.sbs_from_mesh(v, bad, lod=32)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asking for more exits with a message (before any GPU work)."""
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_oracle_baseline_build_computes_the_same_fp32_results():
    """bench.py times the oracle's -O3 / AVX2 build; with FP contraction off it
    must render exactly what the checker build renders."""
    import oracle
    from helpers import preset
    sbs = L.displaced_icosphere(3, 0.8, 0.05, 3, seed=4, lod=16, brick_size=4, brick_pad=0, bytes_per_value=1)
    pairs, _ = L.build_bvh(sbs)
    wv, pr = L.orbit_camera(3, 1)
    p = preset(L.SDF_OCTREE_NODE_INTERSECT_NEWTON)
    a = oracle.Scene(sbs, pairs).render(wv, pr, p, 64, 64, precision=32)
    b = oracle.Scene(sbs, pairs, baseline=True).render(wv, pr, p, 64, 64, precision=32)
    assert (a["prim"] != 0xFFFFFFFF).sum() > 200
    for k in ("t", "n", "prim"):
        assert np.array_equal(a[k], b[k]), k


def test_counter_reduction_averages_dispatches_and_calibrates(tmp_path):
    """bench.py --reduce-counters: per-dispatch averages of the named kernel's
    counters, FETCH/WRITE_SIZE scaled by the copy kernel's reading of a known
    byte count, and the derived L2 hit rate and lane utilisation."""
    import csv
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cols = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    rows = [(1, "__amd_rocclr_copyBuffer", "FETCH_SIZE", 512.0), (1, "__amd_rocclr_copyBuffer", "WRITE_SIZE", 1024.0),
            (2, "void lrtd::k_render<false, false, 2, false>(x)", "FETCH_SIZE", 10.0),
            (2, "void lrtd::k_render<false, false, 2, false>(x)", "WRITE_SIZE", 30.0),
            (3, "void lrtd::k_render<false, false, 2, false>(x)", "FETCH_SIZE", 20.0),
            (3, "void lrtd::k_render<false, false, 2, false>(x)", "WRITE_SIZE", 50.0),
            (2, "void lrtd::k_render<false, false, 2, false>(x)", "TCC_HIT_sum", 3.0),
            (2, "void lrtd::k_render<false, false, 2, false>(x)", "TCC_MISS_sum", 1.0),
            (2, "void lrtd::k_render<false, false, 2, false>(x)", "SQ_THREAD_CYCLES_VALU", 32.0),
            (2, "void lrtd::k_render<false, false, 2, false>(x)", "SQ_ACTIVE_INST_VALU", 1.0),
            (4, "lrtd::k_zero_u32(unsigned int*, unsigned int)", "FETCH_SIZE", 99.0)]
    d = tmp_path / "pass"
    d.mkdir()
    with open(d / "pmc_counter_collection.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(cols)
        w.writerows(rows)
    out = tmp_path / "out.json"
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--reduce-counters", str(d), "--calib-copy-mib", "1",
                        "--out", str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    import json
    j = json.load(open(out))
    # the 1 MiB copy read 1024 KiB while FETCH_SIZE counted 512: scale 2; writes scale 1
    assert j["calibration"]["fetch_scale"] == 2.0 and j["calibration"]["write_scale"] == 1.0
    assert j["hbm_read_bytes"] == int(15.0 * 2 * 1024) and j["hbm_write_bytes"] == int(40.0 * 1024)
    assert j["l2_hit_rate"] == 0.75 and j["valu_lane_utilisation"] == 0.5


def test_fastdiv_matches_integer_division(tmp_path):
    """The render kernels' divisions by launch constants (FastDiv,
    litert_amd/csrc/kernels/fastdiv.hpp) against the host compiler's integer
    division, compiled from the same header with g++."""
    exe = str(tmp_path / "test_fastdiv")
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cpp", "test_fastdiv.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout
