"""CPU-only checks of the product library: it loads, exports every symbol the
C header declares, its host helpers reproduce the reference's known answers,
and argument validation fails cleanly before any GPU work.  The C++ drop-in
headers compile."""
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "umappp_amd.h").read_text()
    return sorted(set(re.findall(r"\b(umx_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(umx):
    L = umx.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", str(umx.library_path())], capture_output=True, text=True,
                        check=True).stdout
    exported = set(re.findall(r" T (umx_\w+)", nm))
    assert set(names) <= exported


def test_library_is_gfx950_code_object(umx):
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", str(umx.library_path())], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    assert "gfx950" in out.stdout


def test_find_ab_known_answers(umx):
    # tests/src/find_ab.cpp:9-19 (values from uwot:::find_ab_params)
    for spread, md, ea, eb in [(1, 0.05, 1.7502250, 0.8420554), (1, 0.01, 1.8956059, 0.800637),
                               (2, 0.1, 0.5446606, 0.8420553)]:
        a, b = umx.find_ab(spread, md)
        assert abs(a - ea) < 1e-4 and abs(b - eb) < 1e-4
    # damping (:21-32) and forced underflow (:34-37)
    a, b = umx.find_ab(200, 1)
    assert abs(a - 0.0004176367) < 1e-6 and abs(b - 0.7955526861) < 1e-4
    a, _ = umx.find_ab(20, 1000)
    assert a < 1e-100


def test_find_ab_fit_is_least_squares_stationary(umx):
    # At the returned (a, b) the least-squares gradient vanishes (to the
    # stopping tolerance): central differences of the sum of squares.
    spread, md = 1.0, 0.1
    a, b = umx.find_ab(spread, md)
    x = np.arange(1, 301) * spread * 3 / 300
    y = np.where(x <= md, 1.0, np.exp(-(x - md) / spread))

    def ss(A, B):
        return np.sum((y - 1 / (1 + A * x ** (2 * B))) ** 2)

    h = 1e-6
    ga = (ss(a + h, b) - ss(a - h, b)) / (2 * h)
    gb = (ss(a, b + h) - ss(a, b - h)) / (2 * h)
    assert abs(ga) < 1e-3 and abs(gb) < 1e-3


def test_choose_num_epochs_known_answers(umx):
    # tests/src/umappp.cpp:315-321
    assert umx.choose_num_epochs(None, 1000) == 500
    assert umx.choose_num_epochs(None, 20000) < 500
    assert umx.choose_num_epochs(None, 10000000) == 201
    assert umx.choose_num_epochs(1000, 1000) == 1000
    assert umx.choose_num_epochs(1000, 20000) == 1000
    assert umx.choose_num_epochs(None, 1000000) == 203
    assert umx.choose_num_epochs(None, 100000) == 230


def test_argument_validation_before_gpu(umx):
    import ctypes as C

    L = umx.lib()
    h = C.c_void_p()
    rp = np.array([0, 1], dtype=np.int64)
    assert L.umx_graph_create(7, 1, rp.ctypes.data_as(C.c_void_p), None, rp.ctypes.data_as(C.c_void_p), 0,
                              C.byref(h)) == -1
    assert b"dtype" in L.umx_last_error()
    assert L.umx_graph_create(0, -1, rp.ctypes.data_as(C.c_void_p), None, None, 0, C.byref(h)) == -5
    assert L.umx_layout_run(None, 5) == -1
    out = C.c_double()
    assert L.umx_find_ab(1.0, 0.1, None, C.byref(out)) == -1


def test_cpp_dropin_headers_compile(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "t.cpp"
    src.write_text(
        '#include "umappp/umappp.hpp"\n'
        '#include "umappp/neighbor_similarities.hpp"\n'
        '#include "umappp/combine_neighbor_sets.hpp"\n'
        "int main() {\n"
        "  umappp::NeighborList<int, float> x(3);\n"
        "  umappp::NeighborSimilaritiesOptions<float> o;\n"
        "  umappp::neighbor_similarities(x, o);\n"
        "  umappp::combine_neighbor_sets(x, 1.0f);\n"
        "  std::vector<float> emb(6);\n"
        "  umappp::Options opt; opt.initialize_method = umappp::InitializeMethod::NONE;\n"
        "  auto st = umappp::initialize(x, 2, emb.data(), opt);\n"
        "  st.run(emb.data(), 10);\n"
        "  std::vector<float> data(30);\n"
        "  auto st2 = umappp::initialize(10, 3, data.data(), 2, emb.data(), opt);\n"
        "  return st.epoch() + st2.epoch();\n"
        "}\n")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-pedantic",
                    f"-I{ROOT / 'include'}", str(src)], check=True)
